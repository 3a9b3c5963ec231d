"""The reference's on-disk format: pickles of tncontract objects.

The reference has no save/load functions of its own; its objects are
pickled whole (``tncontract/tests/random_10site_mps.dat`` is a protocol-3
pickle of a ``MatrixProductState``, ``random_10site_mps_py2.dat`` a
protocol-2 one written by Python 2).  An object's pickle is its ``__dict__``:
a ``Tensor`` is ``{"data": ndarray, "_labels": [str, ...]}``, a
``MatrixProductState`` is ``{"left_label", "right_label", "data": object
ndarray of Tensors, "phys_label"}`` (``MatrixProductStateCanonical`` the
same, ``MatrixProductOperator`` with ``physout_label``/``physin_label``).

``load``/``loads`` read such files WITHOUT unpickling: the opcode stream is
walked by a small stack machine that only records structure (GLOBAL pushes an
inert name, REDUCE/NEWOBJ/OBJ/BUILD build inert records); nothing named in
the file is imported or called, and any class outside the list below is an
error.  Arrays are rebuilt from their raw bytes with ``np.frombuffer`` and
uploaded to the GPU.  ``dump``/``dumps`` write the same layout (protocol 3,
``numpy.core.multiarray._reconstruct`` arrays) so that the reference's own
``pickle.load`` reads what this package writes.
"""
from __future__ import annotations

import io as _io
import pickletools
import struct

import numpy as np

__all__ = ["load", "loads", "dump", "dumps", "read_host", "write_host"]

_T = ("tncontract.tensor", "Tensor")
_CORE = "tncontract.onedim.onedim_core"
_ONE_D = {"MatrixProductState": ("left_label", "right_label", "data", "phys_label"),
          "MatrixProductStateCanonical": ("left_label", "right_label", "data", "phys_label"),
          "MatrixProductOperator": ("left_label", "right_label", "data", "physout_label",
                                    "physin_label"),
          "OneDimensionalTensorNetwork": ("left_label", "right_label", "data")}
_RECON = {("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct")}
_ALLOWED = ({_T, ("numpy", "ndarray"), ("numpy", "dtype"), ("tncontract.label", "Label")}
            | _RECON | {(_CORE, k) for k in _ONE_D})

_MARK = object()


class _Global(tuple):
    pass


class _Obj:
    __slots__ = ("cls", "args", "state")

    def __init__(self, cls, args):
        self.cls, self.args, self.state = cls, args, None


def _key(k):
    """Attribute names: Python 2 pickles store them as byte strings."""
    return k.decode("latin1") if isinstance(k, bytes) else k


def _walk(data):
    """Inert structure of a pickle (no object from the file is created)."""
    stack, memo = [], {}

    def pop_mark():
        items = []
        while True:
            x = stack.pop()
            if x is _MARK:
                return items[::-1]
            items.append(x)

    for op, arg, _ in pickletools.genops(data):
        n = op.name
        if n in ("PROTO", "FRAME"):
            continue
        if n in ("GLOBAL", "STACK_GLOBAL"):
            if n == "GLOBAL":
                mod, name = arg.split(" ", 1)
            else:
                name = stack.pop()
                mod = stack.pop()
            if (mod, name) not in _ALLOWED:
                raise ValueError("unsupported class in pickle: %s.%s" % (mod, name))
            stack.append(_Global((mod, name)))
        elif n in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[arg] = stack[-1]
        elif n == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif n in ("BINGET", "LONG_BINGET", "GET"):
            stack.append(memo[arg])
        elif n == "MARK":
            stack.append(_MARK)
        elif n in ("BININT1", "BININT", "BININT2", "INT", "LONG1", "LONG4", "BINFLOAT",
                   "BINUNICODE", "SHORT_BINUNICODE", "BINUNICODE8", "UNICODE"):
            stack.append(arg)
        elif n in ("SHORT_BINBYTES", "BINBYTES", "BINBYTES8"):
            stack.append(bytes(arg))
        elif n in ("SHORT_BINSTRING", "BINSTRING", "STRING"):
            # Python 2 str: text for labels, raw bytes for array buffers
            stack.append(arg.encode("latin1") if isinstance(arg, str) else arg)
        elif n == "NONE":
            stack.append(None)
        elif n == "NEWFALSE":
            stack.append(False)
        elif n == "NEWTRUE":
            stack.append(True)
        elif n == "EMPTY_TUPLE":
            stack.append(())
        elif n in ("TUPLE1", "TUPLE2", "TUPLE3"):
            k = int(n[-1])
            items = stack[-k:]
            del stack[-k:]
            stack.append(tuple(items))
        elif n == "TUPLE":
            stack.append(tuple(pop_mark()))
        elif n == "EMPTY_LIST":
            stack.append([])
        elif n == "LIST":
            stack.append(pop_mark())
        elif n == "EMPTY_DICT":
            stack.append({})
        elif n == "DICT":
            items = pop_mark()
            stack.append({_key(items[k]): items[k + 1] for k in range(0, len(items), 2)})
        elif n == "APPENDS":
            items = pop_mark()
            stack[-1].extend(items)
        elif n == "APPEND":
            x = stack.pop()
            stack[-1].append(x)
        elif n == "SETITEMS":
            items = pop_mark()
            for k in range(0, len(items), 2):
                stack[-1][_key(items[k])] = items[k + 1]
        elif n == "SETITEM":
            v = stack.pop()
            k = stack.pop()
            stack[-1][_key(k)] = v
        elif n == "REDUCE":
            args = stack.pop()
            fn = stack.pop()
            stack.append(_Obj(fn, args))
        elif n == "NEWOBJ":
            args = stack.pop()
            cls = stack.pop()
            stack.append(_Obj(cls, args))
        elif n == "OBJ":
            items = pop_mark()
            stack.append(_Obj(items[0], tuple(items[1:])))
        elif n == "BUILD":
            st = stack.pop()
            stack[-1].state = st
        elif n == "STOP":
            return stack.pop()
        else:
            raise ValueError("unsupported pickle opcode " + n)
    raise ValueError("truncated pickle (no STOP)")


def _text(x):
    if isinstance(x, bytes):
        return x.decode("latin1")
    if isinstance(x, _Obj) and x.cls == ("tncontract.label", "Label"):
        return _text(x.args[0])
    return str(x)


def _array(obj):
    """Host ndarray (or list of element records for object arrays)."""
    if not (isinstance(obj, _Obj) and obj.cls in _RECON):
        raise ValueError("expected a numpy array record")
    _ver, shape, dtype, fortran, raw = obj.state
    code = _text(dtype.args[0])
    if code == "O8":
        return list(raw), tuple(shape)
    order = _text(dtype.state[1]) if dtype.state is not None else "="
    dt = np.dtype(code).newbyteorder(order if order in "<>" else "=")
    arr = np.frombuffer(raw if isinstance(raw, bytes) else raw.encode("latin1"), dtype=dt)
    return arr.reshape(shape, order="F" if fortran else "C").astype(dt.newbyteorder("="))


def _host_tensor(obj):
    if not (isinstance(obj, _Obj) and obj.cls == _T):
        raise ValueError("expected a tncontract.tensor.Tensor record")
    st = obj.state
    labels = st.get("_labels", st.get("labels"))
    return _array(st["data"]), [_text(x) for x in labels]


def read_host(data):
    """Parse a pickled reference Tensor / 1D network from bytes into host
    data: ("Tensor", array, labels) or (class name, attributes dict with
    "tensors": [(array, labels), ...])."""
    root = _walk(data)
    if not isinstance(root, _Obj):
        raise ValueError("pickle does not hold a tncontract object")
    if root.cls == _T:
        arr, labels = _host_tensor(root)
        return ("Tensor", arr, labels)
    mod, name = root.cls
    if mod != _CORE or name not in _ONE_D:
        raise ValueError("unsupported class %s.%s" % (mod, name))
    st = root.state
    sites, _shape = _array(st["data"])
    attrs = {k: _text(st[k]) for k in _ONE_D[name] if k != "data"}
    attrs["tensors"] = [_host_tensor(s) for s in sites]
    return (name, attrs)


def loads(data):
    """A pickled reference Tensor / MatrixProductState / MatrixProductState-
    Canonical / MatrixProductOperator as the corresponding GPU object."""
    from tncontract_amd.tensor import Tensor
    from tncontract_amd import onedim as od
    h = read_host(data)
    if h[0] == "Tensor":
        return Tensor(np.ascontiguousarray(h[1]), h[2])
    name, attrs = h
    tensors = [Tensor(np.ascontiguousarray(a), lab) for a, lab in attrs.pop("tensors")]
    cls = {"MatrixProductState": od.MatrixProductState,
           "MatrixProductStateCanonical": od.MatrixProductStateCanonical,
           "MatrixProductOperator": od.MatrixProductOperator,
           "OneDimensionalTensorNetwork": od.OneDimensionalTensorNetwork}[name]
    return cls(tensors, **attrs)


def load(path):
    with open(path, "rb") as f:
        return loads(f.read())


# ---------------------------------------------------------------- writer
class _Writer:
    def __init__(self):
        self.b = _io.BytesIO()

    def op(self, code):
        self.b.write(code)

    def glob(self, mod, name):
        self.b.write(b"c" + mod.encode() + b"\n" + name.encode() + b"\n")

    def text(self, s):
        e = s.encode("utf-8")
        self.b.write(b"X" + struct.pack("<I", len(e)) + e)

    def int(self, v):
        if 0 <= v < 256:
            self.b.write(b"K" + struct.pack("<B", v))
        else:
            self.b.write(b"J" + struct.pack("<i", v))

    def raw(self, data):
        if len(data) < 256:
            self.b.write(b"C" + struct.pack("<B", len(data)) + data)
        else:
            self.b.write(b"B" + struct.pack("<I", len(data)) + data)

    def dtype(self, code, order, flags):
        self.glob("numpy", "dtype")
        self.text(code)
        self.op(b"\x89\x88\x87R(")              # False, True, TUPLE3, REDUCE, MARK
        self.int(3)
        self.text(order)
        self.op(b"NNN")
        self.int(-1)
        self.int(-1)
        self.int(flags)
        self.op(b"tb")                          # TUPLE, BUILD

    def array_head(self, shape):
        self.glob("numpy.core.multiarray", "_reconstruct")
        self.glob("numpy", "ndarray")
        self.int(0)
        self.op(b"\x85")                        # TUPLE1
        self.raw(b"b")
        self.op(b"\x87R(")                      # TUPLE3, REDUCE, MARK
        self.int(1)
        self.op(b"(")
        for d in shape:
            self.int(int(d))
        self.op(b"t")

    def ndarray(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.array_head(a.shape)
        self.dtype("f8", "<", 0)
        self.op(b"\x89")                        # not Fortran order
        self.raw(a.astype("<f8").tobytes())
        self.op(b"tb")

    def tensor(self, arr, labels):
        self.glob("tncontract.tensor", "Tensor")
        self.op(b")\x81}(")                     # EMPTY_TUPLE, NEWOBJ, EMPTY_DICT, MARK
        self.text("data")
        self.ndarray(arr)
        self.text("_labels")
        self.op(b"](")
        for x in labels:
            self.text(str(x))
        self.op(b"e")                           # APPENDS
        self.op(b"ub")                          # SETITEMS, BUILD

    def network(self, name, attrs, tensors):
        self.glob(_CORE, name)
        self.op(b")\x81}(")
        for k in _ONE_D[name]:
            self.text(k)
            if k == "data":
                self.array_head((len(tensors),))
                self.dtype("O8", "|", 63)
                self.op(b"\x89](")               # not Fortran, EMPTY_LIST, MARK
                for arr, labels in tensors:
                    self.tensor(arr, labels)
                self.op(b"etb")                  # APPENDS, TUPLE, BUILD
            else:
                self.text(attrs[k])
        self.op(b"ub")


def write_host(kind, *args):
    """Pickle bytes from host data: write_host("Tensor", array, labels) or
    write_host(class name, attributes dict with "tensors": [(array, labels),
    ...]) -- the inverse of read_host."""
    w = _Writer()
    w.op(b"\x80\x03")
    if kind == "Tensor":
        w.tensor(args[0], args[1])
    else:
        if kind not in _ONE_D:
            raise TypeError("cannot write %s in the reference format" % kind)
        attrs = dict(args[0])
        w.network(kind, attrs, attrs["tensors"])
    w.op(b".")
    return w.b.getvalue()


def dumps(obj):
    """Bytes of a protocol-3 pickle in the reference's layout for a Tensor,
    MatrixProductState, MatrixProductStateCanonical, MatrixProductOperator or
    OneDimensionalTensorNetwork of this package (float64 data)."""
    from tncontract_amd.tensor import Tensor
    from tncontract_amd import onedim as od
    if isinstance(obj, Tensor):
        return write_host("Tensor", obj.numpy(), list(obj.labels))
    for name, cls in (("MatrixProductStateCanonical", od.MatrixProductStateCanonical),
                      ("MatrixProductState", od.MatrixProductState),
                      ("MatrixProductOperator", od.MatrixProductOperator),
                      ("OneDimensionalTensorNetwork", od.OneDimensionalTensorNetwork)):
        if isinstance(obj, cls):
            break
    else:
        raise TypeError("cannot write %s in the reference format" % type(obj).__name__)
    attrs = {k: getattr(obj, k) for k in _ONE_D[name] if k != "data"}
    attrs["tensors"] = [(t.numpy(), list(t.labels)) for t in obj]
    return write_host(name, attrs)


def dump(obj, path):
    with open(path, "wb") as f:
        f.write(dumps(obj))
