"""Network contraction ``con`` (tncontract/tncon.py:4-159).

Internal (same-tensor) pairs are traced first; pairs between two tensors are
contracted in argument order, all pairs joining the same two tensors in one
contraction, the result stored at the lower tensor index; finally the tensor
product of the disconnected components is returned.  Each pairwise step is
one MFMA contraction on the GPU.
"""
from tncontract_amd import tensor as tn


def con(*args):
    tensor_list, contract_list = [], []
    for x in args:
        if isinstance(x, list):
            if isinstance(x[0], tn.Tensor):
                tensor_list.extend(x)
            else:
                contract_list.extend(x)
        elif isinstance(x, tn.Tensor):
            tensor_list.append(x)
        else:
            contract_list.append(x)
    tensor_list = [t.copy() for t in tensor_list]
    flat = [lab for pair in contract_list for lab in pair]
    if len(set(flat)) != len(flat):
        raise ValueError("Index found in more than one contraction pair.")
    owner = {}
    for i, t in enumerate(tensor_list):
        for lab in t.labels:
            if lab in owner:
                raise ValueError("Index label " + lab + " found in two tensors."
                                 + " Tensors must have unique index labelling.")
            owner[lab] = i
    internal, pairwise, keys = [], [], []
    for c in contract_list:
        a, b = owner[c[0]], owner[c[1]]
        if a == b:
            internal.append(c)
            continue
        key = tuple(sorted((a, b)))
        if key in keys:
            k = keys.index((a, b))
            if not isinstance(pairwise[k][0], list):
                pairwise[k][0] = [pairwise[k][0]]
                pairwise[k][1] = [pairwise[k][1]]
            pairwise[k][0].append(c[0])
            pairwise[k][1].append(c[1])
        else:
            pairwise.append(list(c))
            keys.append(key)
    for c in internal:
        tensor_list[owner[c[0]]].trace(c[0], c[1])
    component = list(range(len(tensor_list)))
    for c in pairwise:
        first0 = c[0][0] if isinstance(c[0], list) else c[0]
        first1 = c[1][0] if isinstance(c[1], list) else c[1]
        d, e = owner[first0], owner[first1]
        if d == e:
            tensor_list[d].trace(c[0], c[1])
        else:
            if d < e:
                tensor_list[d] = tn.contract(tensor_list[d], tensor_list[e], c[0], c[1])
                component[e] = d
            else:
                tensor_list[e] = tn.contract(tensor_list[e], tensor_list[d], c[1], c[0])
                component[d] = e
            lo = min(d, e)
            for lab in tensor_list[lo].labels:
                owner[lab] = lo
    return tn.tensor_product(*[tensor_list[component.index(x)] for x in set(component)])
