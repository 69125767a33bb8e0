"""Synthetic loops (include/op2amd_synth.h) and the MLP cost model
(include/op2amd_costmodel.h)."""
import ctypes as C
import math

import numpy as np

from ._lib import lib
from .api import PLAN_ATOMIC, PLAN_COLOR, PLAN_STAGED, PLAN_GATHER, PLAN_OWNER
from .mesh import ORDERINGS

DIRECT = 0
OP_READ, OP_INC, OP_RW = 0, 3, 2
ACCESS = {"read": OP_READ, "inc": OP_INC, "rw": OP_RW}


class Desc(C.Structure):
    _fields_ = [("I", C.c_int), ("J", C.c_int), ("ordering", C.c_int), ("arity", C.c_int),
                ("dim", C.c_int), ("access", C.c_int), ("is_double", C.c_int),
                ("k_fma", C.c_int), ("seed", C.c_uint64)]


class Shape(C.Structure):
    _fields_ = [("set_size", C.c_int), ("target_size", C.c_int),
                ("mean_bandwidth", C.c_double)]


lib.op2amd_synth_time.restype = C.c_int
lib.op2amd_synth_time.argtypes = [C.POINTER(Desc), C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(Shape)]
lib.op2amd_synth_run_once.restype = C.c_int
lib.op2amd_synth_run_once.argtypes = [C.POINTER(Desc), C.c_int, C.c_int,
                                      C.POINTER(C.c_double), C.c_int]
lib.op2amd_mlp_create.restype = C.c_void_p
lib.op2amd_mlp_create.argtypes = [C.c_int, C.c_int, C.c_uint64]
lib.op2amd_mlp_destroy.argtypes = [C.c_void_p]
lib.op2amd_mlp_num_params.restype = C.c_int
lib.op2amd_mlp_num_params.argtypes = [C.c_void_p]
for fn in ("op2amd_mlp_get_params", "op2amd_mlp_set_params"):
    getattr(lib, fn).restype = C.c_int
    getattr(lib, fn).argtypes = [C.c_void_p, C.POINTER(C.c_float)]
lib.op2amd_mlp_forward.restype = C.c_int
lib.op2amd_mlp_forward.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int,
                                   C.POINTER(C.c_float)]
lib.op2amd_mlp_loss_grad.restype = C.c_int
lib.op2amd_mlp_loss_grad.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
lib.op2amd_mlp_set_weight_decay.restype = C.c_int
lib.op2amd_mlp_set_weight_decay.argtypes = [C.c_void_p, C.c_float]
lib.op2amd_mlp_train.restype = C.c_int
lib.op2amd_mlp_train.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]


_INSTALLED = None
lib.op2h_set_plan_model.restype = C.c_int
lib.op2h_set_plan_model.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.c_int]
lib.op2h_set_loop_arith.argtypes = [C.c_char_p, C.c_int]
lib.op2h_plan_model_loop.restype = C.c_int
lib.op2h_plan_model_loop.argtypes = [C.c_char_p, C.POINTER(C.c_double)]


def model_loop(name):
    """The runtime's description of loop `name` at its model-driven first
    call: dict of set_size, target_size, arity, dim, is_double, k,
    mean_bandwidth, access (None if the loop has none)."""
    d = (C.c_double * 8)()
    if lib.op2h_plan_model_loop(name.encode(), d):
        return None
    return {"set_size": d[0], "target_size": d[1], "arity": int(d[2]), "dim": int(d[3]),
            "is_double": int(d[4]), "k": d[5], "mean_bandwidth": d[6],
            "access": ("read", "inc", "rw")[int(d[7])]}


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def desc_from(n_target, arity, dim, access, is_double, ordering, k, seed):
    """Descriptor whose loop set has about n_target elements: edges (about
    2 I J) for arity 2, cells (I J) otherwise; I = 2 J."""
    cells = n_target / 2 if arity == 2 else n_target
    J = max(4, int(round(math.sqrt(cells / 2))))
    return Desc(2 * J, J, ORDERINGS[ordering], arity, dim, ACCESS[access], int(is_double),
                k, seed)


BLOCKS = (64, 128, 256, 512, 1024)


def variants(access):
    """Candidate (plan, block) pairs of the exhaustive sweep for an access mode."""
    blocks = BLOCKS
    if access == "read":
        return [(DIRECT, b) for b in blocks]
    plans = [PLAN_COLOR, PLAN_GATHER] + ([PLAN_ATOMIC, PLAN_OWNER] if access == "inc" else [])
    out = [(p, b) for p in plans for b in blocks]
    out += [(PLAN_STAGED, b) for b in (128, 256, 512)]
    return out


def time_variants(desc, var, reps=3):
    n = len(var)
    plan = (C.c_int * n)(*[p for p, _ in var])
    block = (C.c_int * n)(*[b for _, b in var])
    ms = (C.c_double * n)()
    shape = Shape()
    if lib.op2amd_synth_time(C.byref(desc), n, plan, block, reps, ms, C.byref(shape)):
        raise RuntimeError("synthetic loop timing failed")
    return list(ms), {"set_size": shape.set_size, "target_size": shape.target_size,
                      "mean_bandwidth": shape.mean_bandwidth}


lib.op2amd_synth_check.restype = C.c_int
lib.op2amd_synth_check.argtypes = [C.POINTER(Desc), C.c_void_p, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.POINTER(C.c_longlong)]


def desc_mesh(desc):
    """The generator mesh a descriptor's loop runs over."""
    from .mesh import generate, ORDERINGS as _O
    name = {v: k for k, v in _O.items()}[desc.ordering]
    return generate(desc.I, desc.J, desc.seed, name)


def check_plans(desc, var, ref=(PLAN_GATHER, 128), mesh=None):
    """Run the reference variant and each (plan, block) of `var` once from the
    same seeded values; per variant (max |w - w_ref| / max |w_ref|, elements
    whose bits differ), (-1, -1) where the variant does not apply.  `mesh`:
    the descriptor's mesh if already generated (desc_mesh)."""
    n = len(var)
    plan = (C.c_int * n)(*[p for p, _ in var])
    block = (C.c_int * n)(*[b for _, b in var])
    dev = (C.c_double * n)()
    mis = (C.c_longlong * n)()
    mp = C.byref(mesh.c()) if mesh is not None else None
    rc = lib.op2amd_synth_check(C.byref(desc), mp, ref[0], ref[1], n, plan, block, dev, mis)
    if rc:
        raise RuntimeError(f"synthetic plan check failed ({rc})")
    return list(zip(list(dev), list(mis)))


def run_once(desc, plan, block, size):
    out = np.zeros(size)
    rc = lib.op2amd_synth_run_once(C.byref(desc), plan, block,
                                   out.ctypes.data_as(C.POINTER(C.c_double)), size)
    if rc:
        raise RuntimeError(f"synthetic loop run failed ({rc})")
    return out


class MLP:
    """The library's HIP MLP (nin -> hidden -> hidden -> 1, ReLU, FP32)."""

    def __init__(self, nin, hidden=64, seed=0):
        self.h = lib.op2amd_mlp_create(nin, hidden, seed)
        if not self.h:
            raise RuntimeError("MLP creation failed (no GPU?)")
        self.nin, self.hidden = nin, hidden
        self.np = lib.op2amd_mlp_num_params(self.h)

    def params(self):
        p = np.zeros(self.np, np.float32)
        lib.op2amd_mlp_get_params(self.h, _fp(p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        lib.op2amd_mlp_set_params(self.h, _fp(p))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[0], np.float32)
        lib.op2amd_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        g = np.zeros(self.np, np.float32)
        lib.op2amd_mlp_loss_grad(self.h, _fp(x), _fp(t), x.shape[0], C.byref(loss), _fp(g))
        return loss.value, g

    def set_weight_decay(self, wd):
        if lib.op2amd_mlp_set_weight_decay(self.h, wd):
            raise ValueError("weight decay must be >= 0")

    def train(self, x, t, epochs, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        lib.op2amd_mlp_train(self.h, _fp(x), _fp(t), x.shape[0], epochs, lr, C.byref(loss))
        return loss.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.op2amd_mlp_destroy(self.h)
            self.h = None


ORDER_NAMES = ("random", "rcm", "hilbert")
PLAN_IDS = (DIRECT, PLAN_ATOMIC, PLAN_COLOR, PLAN_STAGED, PLAN_GATHER, PLAN_OWNER)


def features(rec, plan, block):
    """Feature vector of one (loop, variant): sizes, shape, arithmetic and the
    variant itself.  Everything in it is known for a real loop at its first
    call (the numbering enters only through the map's mean bandwidth), so the
    runtime builds the same vector (op2h_set_plan_model)."""
    f = [math.log10(max(1, rec["set_size"])), math.log10(max(1, rec["target_size"])),
         rec["arity"] / 4.0, math.log2(rec["dim"]) / 3.0, float(rec["is_double"]),
         rec["k"] / 64.0, math.log10(1.0 + rec["mean_bandwidth"]) / 7.0,
         float(rec["access"] == "read"), float(rec["access"] == "inc"),
         float(rec["access"] == "rw")]
    f += [float(plan == p) for p in PLAN_IDS]
    # block size both as a magnitude and one-hot: its effect is not monotonic
    f += [math.log2(block) / 10.0]
    f += [float(block == b) for b in BLOCKS]
    return f


# ---------------------------------------------------------------------------
# Plan selection with the cost model
# ---------------------------------------------------------------------------

def sweep(descs, reps=5, log=None):
    """Exhaustive timing of every variant of every loop description
    (dicts with keys n, arity, dim, access, is_double, ordering, k, seed)."""
    recs = []
    for i, p in enumerate(descs):
        d = desc_from(p["n"], p["arity"], p["dim"], p["access"], p["is_double"],
                      p["ordering"], p["k"], p["seed"])
        var = variants(p["access"])
        ms, shape = time_variants(d, var, reps=reps)
        rec = dict(p)
        rec.update(variants=[list(v) for v in var], ms=ms, **shape)
        recs.append(rec)
        if log:
            best = min(t for t in ms if t > 0)
            log(f"  loop {i}: n={shape['set_size']} a={p['arity']} d={p['dim']} {p['access']} "
                f"{'f64' if p['is_double'] else 'f32'} {p['ordering']} k={p['k']} "
                f"best {best * 1e3:.1f}us")
    return recs


def random_descs(count, seed, n_lo=1e4, n_hi=1e7):
    """Seeded loop descriptions: set sizes log-uniform in [n_lo, n_hi]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        out.append(dict(n=float(10 ** rng.uniform(math.log10(n_lo), math.log10(n_hi))),
                        arity=int(rng.integers(1, 5)), dim=int(rng.choice([1, 2, 4, 8])),
                        access=str(rng.choice(["read", "inc", "rw"])),
                        is_double=int(rng.integers(0, 2)),
                        ordering=str(rng.choice(list(ORDER_NAMES))),
                        k=int(rng.integers(4, 65)), seed=100 + i + 1000 * seed))
    return out


def valid_variants(rec):
    return [(tuple(v), t) for v, t in zip(rec["variants"], rec["ms"]) if t > 0]


def regret(recs, ids, choose):
    """time(chosen) / time(fastest) - 1 per loop."""
    out = []
    for i in ids:
        valid = valid_variants(recs[i])
        best = min(t for _, t in valid)
        out.append(dict(valid)[choose(recs[i])] / best - 1.0)
    return np.array(out)


def default_choice(rec):
    """The runtime's static choice without a model: the reference's
    OP_STAGE_ALL plan (staged, block 256) for increments and read-write
    loops, block 256 for conflict-free loops."""
    return (DIRECT, 256) if rec["access"] == "read" else (PLAN_STAGED, 256)


def best_fixed_choice(recs, ids):
    """The single (plan, block) per access mode with the lowest mean regret
    over the loops `ids`: the best a model-free static rule can do."""
    table = {}
    for acc in ("read", "inc", "rw"):
        mine = [i for i in ids if recs[i]["access"] == acc]
        if not mine:
            continue
        cands = [tuple(v) for v in recs[mine[0]]["variants"]]
        score = {c: np.mean(regret(recs, mine, lambda r, c=c: c)) for c in cands
                 if all(dict(valid_variants(recs[i])).get(c) for i in mine)}
        table[acc] = min(score, key=score.get)
    return lambda rec: table.get(rec["access"], default_choice(rec))


class CostModel:
    """An ensemble of the library's HIP MLPs predicting each variant's log
    time relative to the mean over the loop's variants (only the ranking
    within a loop matters for the choice)."""

    def __init__(self, hidden=64, members=5, seed=0, weight_decay=1e-4, lr=3e-3):
        self.hidden, self.members, self.seed = hidden, members, seed
        self.wd, self.lr = weight_decay, lr
        self.nets = []

    def _rows(self, recs, ids):
        X, y = [], []
        for i in ids:
            r = recs[i]
            valid = valid_variants(r)
            lt = np.log([t for _, t in valid])
            for (v, _), t in zip(valid, lt - lt.mean()):
                X.append(features(r, v[0], v[1]))
                y.append(t)
        return np.array(X, np.float32), np.array(y, np.float32)

    def _start(self, X):
        self.mu, self.sd = X.mean(axis=0), X.std(axis=0) + 1e-6
        self.nets = []
        for m in range(self.members):
            net = MLP(X.shape[1], self.hidden, seed=self.seed * 1000 + m)
            net.set_weight_decay(self.wd)
            self.nets.append(net)

    def _train(self, Xn, y, epochs):
        return [net.train(Xn, y, epochs, self.lr) for net in self.nets]

    def fit(self, recs, ids, epochs):
        X, y = self._rows(recs, ids)
        self._start(X)
        self.loss = self._train((X - self.mu) / self.sd, y, epochs)
        self.epochs = epochs
        return self

    def fit_validated(self, recs, ids, max_epochs=4000, chunk=250, holdout=0.2):
        """Choose the epoch count on a held-out part of `ids`, then fit on all
        of `ids` with it."""
        ids = list(ids)
        rng = np.random.default_rng(self.seed + 7)
        perm = rng.permutation(len(ids))
        nval = max(1, int(len(ids) * holdout))
        val = [ids[j] for j in perm[:nval]]
        tr = [ids[j] for j in perm[nval:]]
        X, y = self._rows(recs, tr)
        self._start(X)
        Xn = (X - self.mu) / self.sd
        curve = []
        for e in range(chunk, max_epochs + 1, chunk):
            self._train(Xn, y, chunk)
            curve.append((e, float(np.mean(regret(recs, val, self.choose)))))
        best_e = min(curve, key=lambda c: (c[1], c[0]))[0]
        self.validation_curve = curve
        return self.fit(recs, ids, best_e)

    def predict(self, X):
        Xn = ((np.asarray(X, np.float32) - self.mu) / self.sd).astype(np.float32)
        return np.mean([net.forward(Xn) for net in self.nets], axis=0)

    def ranked(self, rec):
        """The loop's valid variants, best predicted first."""
        valid = valid_variants(rec)
        pred = self.predict([features(rec, v[0], v[1]) for v, _ in valid])
        return [valid[i][0] for i in np.argsort(pred, kind="stable")]

    def choose(self, rec):
        return self.ranked(rec)[0]

    def install(self, top=1):
        """Make the runtime pick the plan of every loop left on AUTO with this
        model at the loop's first call (op2h_set_plan_model): the best
        predicted variant, or with top > 1 the fastest of the `top` best,
        timed on the loop's next calls.  The model stays installed until
        uninstall()."""
        global _INSTALLED
        handles = (C.c_void_p * len(self.nets))(*[n.h for n in self.nets])
        mu = np.ascontiguousarray(self.mu, np.float32)
        sd = np.ascontiguousarray(self.sd, np.float32)
        if lib.op2h_set_plan_model(handles, len(self.nets), len(mu), _fp(mu), _fp(sd), top):
            raise ValueError("the runtime refused the model (feature count or scale)")
        _INSTALLED = (self, handles, mu, sd)  # keeps the networks alive
        return self

    @staticmethod
    def uninstall():
        global _INSTALLED
        lib.op2h_set_plan_model(None, 0, 0, None, None, 1)
        _INSTALLED = None

    def choose_timed(self, rec, k):
        """Model-pruned tuning: time only the k best predicted variants (here
        read from the sweep) and keep the fastest."""
        times = dict(valid_variants(rec))
        return min(self.ranked(rec)[:k], key=lambda v: times[v])
