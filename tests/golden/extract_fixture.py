"""Extract the reference's pickled 10-site MPS fixture into plain arrays.

The reference's test fixture ``tncontract/tests/random_10site_mps.dat`` is a
protocol-3 pickle of a ``MatrixProductState``.  It is NOT unpickled here:
this script walks the opcode stream with ``pickletools.genops`` and a small
stack machine that only records structure -- GLOBAL pushes a name marker,
REDUCE/NEWOBJ/BUILD build inert tuples, nothing from the file is imported or
called.  The float64 site arrays are then rebuilt with ``np.frombuffer`` from
the raw bytes, and the result is written to ``random_10site_mps.npz`` (arrays
plus a JSON label list).  The .npz is the committed fixture; this script is
how it was made (it reads /root/reference, so it runs only where that exists).

Usage:  python tests/golden/extract_fixture.py
"""
import json
import os
import pickletools

import numpy as np

SRC = "/root/reference/tncontract/tests/random_10site_mps.dat"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "random_10site_mps.npz")

_MARK = object()


class Global(tuple):
    pass


class Call(tuple):
    pass


class Obj:
    def __init__(self, cls, args):
        self.cls, self.args, self.state = cls, args, None


def walk(data):
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
        if n == "PROTO":
            continue
        if n == "GLOBAL":
            mod, name = arg.split(" ", 1)
            stack.append(Global((mod, name)))
        elif n in ("BINPUT", "LONG_BINPUT", "MEMOIZE"):
            memo[arg if n != "MEMOIZE" else len(memo)] = stack[-1]
        elif n in ("BINGET", "LONG_BINGET"):
            stack.append(memo[arg])
        elif n == "MARK":
            stack.append(_MARK)
        elif n in ("BININT1", "BININT", "BININT2", "BINUNICODE", "SHORT_BINUNICODE",
                   "SHORT_BINBYTES", "BINBYTES", "BINFLOAT", "SHORT_BINSTRING", "BINSTRING"):
            stack.append(arg)
        elif n == "NONE":
            stack.append(None)
        elif n == "NEWFALSE":
            stack.append(False)
        elif n == "NEWTRUE":
            stack.append(True)
        elif n == "EMPTY_TUPLE":
            stack.append(())
        elif n == "TUPLE1":
            stack.append((stack.pop(),))
        elif n == "TUPLE2":
            b = stack.pop(); a = stack.pop(); stack.append((a, b))
        elif n == "TUPLE3":
            c = stack.pop(); b = stack.pop(); a = stack.pop(); stack.append((a, b, c))
        elif n == "TUPLE":
            stack.append(tuple(pop_mark()))
        elif n == "EMPTY_LIST":
            stack.append([])
        elif n == "EMPTY_DICT":
            stack.append({})
        elif n == "APPENDS":
            items = pop_mark(); stack[-1].extend(items)
        elif n == "APPEND":
            x = stack.pop(); stack[-1].append(x)
        elif n == "SETITEMS":
            items = pop_mark()
            for k in range(0, len(items), 2):
                stack[-1][items[k]] = items[k + 1]
        elif n == "SETITEM":
            v = stack.pop(); k = stack.pop(); stack[-1][k] = v
        elif n == "REDUCE":
            args = stack.pop(); fn = stack.pop(); stack.append(Obj(fn, args))
        elif n == "NEWOBJ":
            args = stack.pop(); cls = stack.pop(); stack.append(Obj(cls, args))
        elif n == "BUILD":
            st = stack.pop(); stack[-1].state = st
        elif n == "STOP":
            return stack.pop()
        else:
            raise ValueError("unsupported opcode " + n)
    raise ValueError("no STOP")


def ndarray_from(obj):
    """Rebuild an ndarray from the inert (_reconstruct, state) record."""
    assert obj.cls == ("numpy.core.multiarray", "_reconstruct")
    _ver, shape, dtype, fortran, raw = obj.state
    code = dtype.args[0]
    if code == "O8":
        return shape, raw  # object array: raw is the list of element records
    dt = np.dtype(code).newbyteorder(dtype.state[1] if dtype.state[1] in "<>" else "=")
    arr = np.frombuffer(raw if isinstance(raw, bytes) else raw.encode("latin1"), dtype=dt)
    return arr.reshape(shape, order="F" if fortran else "C").astype(np.float64)


def main():
    root = walk(open(SRC, "rb").read())
    assert root.cls == ("tncontract.onedim.onedim_core", "MatrixProductState")
    st = root.state
    shape, sites = ndarray_from(st["data"])
    arrays, labels = {}, []
    for i, site in enumerate(sites):
        s = site.state
        arrays["site%d" % i] = ndarray_from(s["data"])
        labels.append(list(s["_labels"]))
    meta = {"left_label": st["left_label"], "right_label": st["right_label"],
            "phys_label": st["phys_label"], "labels": labels, "source": SRC}
    np.savez(DST, meta=np.array(json.dumps(meta)), **arrays)
    print("wrote", DST, [arrays["site%d" % i].shape for i in range(len(sites))])


if __name__ == "__main__":
    main()
