"""Tensor-network plan execution on the device.

Mirrors tenet-network's `Network::new(inputs, conj, codomain_splits, output,
output_codomain_rank)` / `plan(tensors, optimizer)` / `PlannedNetwork::execute`
(tenet-network/src/network.rs:235-339, :1609-1640, schedule compile
:1960-2115) for the device backend.  The reference's device route executes
only canonical whole-codomain x whole-domain steps and rejects the rest
(`execute_cuda` preflight, network.rs:1467-1606); here every step is a
general contraction (operand recoupling + grouped block GEMM + output
recoupling), so any schedule runs on the GPU.

Labels: `inputs[i]` are operand i's labels in flat leg order (codomain then
domain) of the tensor as passed; a `conj` operand is contracted as its lazy
adjoint, whose legs are the domain legs then the codomain legs, so its labels
rotate by the codomain rank.  Every label is either open (once, and listed in
`output`) or contracted (twice, on two different operands).
"""
from itertools import combinations


class NetworkError(ValueError):
    pass


class GreedyOptimizer:
    """Repeatedly contract the pair sharing a label with the smallest dense
    cost (product of the dimensions of all labels of the pair); ties go to the
    lowest operand positions."""

    def order(self, labels, dims):
        live = {i: list(l) for i, l in enumerate(labels)}
        steps = []
        nxt = len(labels)
        while len(live) > 1:
            best = None
            for a, b in combinations(sorted(live), 2):
                shared = set(live[a]) & set(live[b])
                if not shared:
                    continue
                cost = 1
                for lab in set(live[a]) | set(live[b]):
                    cost *= dims[lab]
                key = (cost, a, b)
                if best is None or key < best:
                    best = key
            if best is None:
                raise NetworkError("network is disconnected (outer products are not supported)")
            _, a, b = best
            steps.append((a, b))
            shared = set(live[a]) & set(live[b])
            live[nxt] = [x for x in live[a] if x not in shared] + [x for x in live[b] if x not in shared]
            del live[a], live[b]
            nxt += 1
        return steps


class LabelOrderOptimizer:
    """Contract in the order the given contracted labels are listed
    (the reference's LabelOrderDenseOptimizer): each label merges the two
    current tensors holding it, together with all labels they share."""

    def __init__(self, order):
        self.labels = list(order)

    def order(self, labels, dims):
        owner = {}
        live = {i: list(l) for i, l in enumerate(labels)}
        for i, l in live.items():
            for x in l:
                owner.setdefault(x, []).append(i)
        steps = []
        nxt = len(labels)
        for lab in self.labels:
            holders = [i for i in live if lab in live[i]]
            if len(holders) != 2:
                continue  # already contracted with an earlier label
            a, b = sorted(holders)
            shared = set(live[a]) & set(live[b])
            steps.append((a, b))
            live[nxt] = [x for x in live[a] if x not in shared] + [x for x in live[b] if x not in shared]
            del live[a], live[b]
            nxt += 1
        if len(live) != 1:
            raise NetworkError("label order does not connect every operand")
        return steps


class Network:
    def __init__(self, inputs, conj, codomain_splits, output, output_codomain_rank=None):
        if not (len(inputs) == len(conj) == len(codomain_splits)):
            raise NetworkError("operand marker lists must match operand count")
        if output_codomain_rank is not None and output_codomain_rank > len(output):
            raise NetworkError("output codomain rank exceeds output rank")
        count = {}
        for i, labs in enumerate(inputs):
            if len(set(labs)) != len(labs):
                raise NetworkError(f"operand {i} repeats a label (traces are not supported)")
            for x in labs:
                count[x] = count.get(x, 0) + 1
        if any(c > 2 for c in count.values()):
            raise NetworkError("a label appears on more than two operands")
        open_labels = {x for x, c in count.items() if c == 1}
        if set(output) != open_labels or len(output) != len(open_labels):
            raise NetworkError("output labels must be exactly the open labels, each once")
        self.inputs = [list(l) for l in inputs]
        self.conj = list(conj)
        self.codomain_splits = list(codomain_splits)
        self.output = list(output)
        self.output_codomain_rank = output_codomain_rank

    def _lowered(self, tensors):
        """Labels of the tensors as contracted (conj operands rotated)."""
        if len(tensors) != len(self.inputs):
            raise NetworkError(f"plan has {len(self.inputs)} operands but {len(tensors)} tensors were given")
        labels, dims = [], {}
        for i, (t, labs) in enumerate(zip(tensors, self.inputs)):
            if t.space.rank != len(labs):
                raise NetworkError(f"operand {i} has rank {t.space.rank}, labels give {len(labs)}")
            split = self.codomain_splits[i]
            if split is not None and split != t.space.nout:
                raise NetworkError(f"operand {i} codomain split {t.space.nout} differs from the written {split}")
            legs = list(t.space.codomain) + list(t.space.domain)
            for lab, leg in zip(labs, legs):
                d = leg.dim()
                if dims.setdefault(lab, d) != d:
                    raise NetworkError(f"label {lab!r} joins legs of different dimensions")
            k = t.space.nout
            labels.append(labs[k:] + labs[:k] if self.conj[i] else list(labs))
        return labels, dims

    def plan(self, tensors, optimizer=None):
        labels, dims = self._lowered(tensors)
        steps = (optimizer or GreedyOptimizer()).order(labels, dims) if len(labels) > 1 else []
        return PlannedNetwork(self, labels, steps)

    def plan_with(self, tensors, steps):
        """Wrap an explicit pairwise order [(slot_a, slot_b), ...]; results take
        slots len(inputs), len(inputs) + 1, ..."""
        labels, _ = self._lowered(tensors)
        return PlannedNetwork(self, labels, list(steps))


class PlannedNetwork:
    """A compiled pairwise schedule: per step the contracted axes; the final
    permutation to the requested output labels and split is folded into the
    last contraction's output permutation."""

    def __init__(self, network, labels, steps):
        self.network = network
        n = len(labels)
        cur = {i: list(l) for i, l in enumerate(labels)}
        self.steps = []
        for si, (a, b) in enumerate(steps):
            if a not in cur or b not in cur:
                raise NetworkError("step uses an operand that is already consumed")
            la, lb = cur.pop(a), cur.pop(b)
            axes_a = [i for i, x in enumerate(la) if x in lb]
            axes_b = [lb.index(la[i]) for i in axes_a]
            if not axes_a:
                raise NetworkError("step contracts no label (outer products are not supported)")
            res = [x for i, x in enumerate(la) if i not in axes_a] + [x for i, x in enumerate(lb) if i not in axes_b]
            self.steps.append({"lhs": a, "rhs": b, "result": n + si, "axes_a": axes_a, "axes_b": axes_b,
                               "labels": res, "nout": len(la) - len(axes_a), "out_perm": None})
            cur[n + si] = res
        if len(cur) != 1:
            raise NetworkError("schedule leaves more than one tensor")
        (self.final_slot, final_labels), = cur.items()
        out = network.output
        split = network.output_codomain_rank if network.output_codomain_rank is not None else len(out)
        perm = [final_labels.index(x) for x in out]
        if self.steps:
            last = self.steps[-1]
            if perm != list(range(len(perm))) or split != last["nout"]:
                last["out_perm"], last["nout"] = perm, split
            self.final_permutation = None
        else:
            self.final_permutation = (perm[:split], perm[split:])

    def execute(self, tensors, precision=None):
        net = self.network
        slots = {}
        for i, t in enumerate(tensors):
            slots[i] = t.adjoint() if net.conj[i] else t
        for st in self.steps:
            lhs, rhs = slots.pop(st["lhs"]), slots.pop(st["rhs"])
            slots[st["result"]] = lhs.contract(st["axes_a"], rhs, st["axes_b"], st["out_perm"], st["nout"],
                                               precision=precision)
        out = slots.pop(self.final_slot)
        if self.final_permutation is not None:
            p1, p2 = self.final_permutation
            if net.conj[0] and out._adjoint:
                raise NetworkError("a single adjoint operand must be composed before it is permuted")
            out = out.permute(p1, p2)
        return out
