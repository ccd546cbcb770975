"""Learned message-passing preconditioner of the Chorin pressure solve.

The network (``zj_mpnn_*`` in the C ABI) is the reference's message-passing
layer (``zoomy_jax/gnn_blueprint/graph_mp_multibranch.py:9-16``) made linear
in the residual: three gated layers of hidden width 64 on the graph of 2x2
cell blocks, BF16 MFMA products with FP32 accumulation.  It is a learned
smoother after the FP32 multigrid V-cycle V (the role of ``smooth_step`` in
the reference's learned V-cycle, ``vcycle_structured_gnn.py:74-122``,
``:145-200``):

    x = V r;    z = x + N(r - S x)

and the preconditioner of FP64 flexible CG.

Training (``train``) follows the reference's recipe (Adam, learning rate of
order 1e-3, target the exact solution u* = A^-1 f,
``train_vcycle_structured_poisson.py:266-312``) with the error of one learned
V-cycle measured in the operator's energy norm,

    loss = (z - p)^T S (z - p) / (p^T rhs),     z = M rhs,

on (rhs, p) pairs that the plain solver generates: the pressure equations
of Chorin runs and noise right-hand sides (``solver_pairs``), and the
right-hand sides S e of random errors e (``error_pairs``), each solved to
1e-8.  The network's input r - S V r does not depend on its parameters, so
the gradient is the network's own backward.
"""
from __future__ import annotations

import ctypes
import pathlib
import time
from typing import Optional

import numpy as np
import torch

from . import _lib as L

HID, NL, NF = 64, 3, 16
SHAPES = [("Win", (4, HID))]
for _l in range(NL):
    SHAPES += [(f"W{_l}", (2 * HID, HID)), (f"U{_l}", (NF, HID))]
SHAPES += [("Wout", (HID, 4))]
DEFAULT_WEIGHTS = pathlib.Path(__file__).resolve().parent / "data" / \
    "mpnn_chorin.npz"


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def init_params(seed: int = 0) -> np.ndarray:
    """Scaled-normal initialisation with W_out = 0: the untrained network is
    the zero map, so the learned V-cycle starts as a pre-smoothed
    coarse-grid correction."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name == "Win":
            a = rng.standard_normal(shape) * 0.5
        elif name.startswith("W") and name != "Wout":
            a = rng.standard_normal(shape) / np.sqrt(shape[0])
        elif name.startswith("U"):
            a = rng.standard_normal(shape) * 0.1
        else:
            a = np.zeros(shape)
        parts.append(a.ravel())
    return np.concatenate(parts).astype(np.float32)


class MPPreconditioner:
    """Device network: FP32 master parameters, BF16 layer weights, Adam
    moments and gradient accumulators."""

    def __init__(self, params: Optional[np.ndarray] = None, seed: int = 0):
        lib = L.lib()
        self.n_params = int(lib.zj_mpnn_param_count())
        h = ctypes.c_void_p()
        L.check(lib.zj_mpnn_create(ctypes.byref(h)), "zj_mpnn_create")
        self.handle = h
        self.set_params(init_params(seed) if params is None else params)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and L is not None and L._lib is not None:
            L._lib.zj_mpnn_destroy(h)
            self.handle = None

    # ── parameters ────────────────────────────────────────────────────
    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        if flat.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters")
        L.check(L.lib().zj_mpnn_set_params(self.handle, flat.ctypes.data,
                                           flat.size), "zj_mpnn_set_params")

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        L.check(L.lib().zj_mpnn_get_params(self.handle, out.ctypes.data,
                                           out.size), "zj_mpnn_get_params")
        return out

    def get_grad(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        L.check(L.lib().zj_mpnn_get_grad(self.handle, out.ctypes.data,
                                         out.size), "zj_mpnn_get_grad")
        return out

    def set_fused(self, on: bool = True):
        """Single-launch inference kernel (same result as layer by layer)."""
        L.check(L.lib().zj_mpnn_set_fused(self.handle, int(bool(on))),
                "zj_mpnn_set_fused")

    def zero_grad(self):
        L.check(L.lib().zj_mpnn_zero_grad(self.handle, _stream()),
                "zj_mpnn_zero_grad")

    def adam(self, lr=3e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        L.check(L.lib().zj_mpnn_adam(self.handle, lr, beta1, beta2, eps,
                                     _stream()), "zj_mpnn_adam")

    def save(self, path):
        np.savez(path, params=self.get_params(), hidden=HID, layers=NL)

    @classmethod
    def load(cls, path=None):
        with np.load(DEFAULT_WEIGHTS if path is None else path) as f:
            if int(f["hidden"]) != HID or int(f["layers"]) != NL:
                raise ValueError("weights of another architecture")
            return cls(params=f["params"])

    # ── network ───────────────────────────────────────────────────────
    @staticmethod
    def _f32(t, shape=None):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("expected a contiguous float32 CUDA tensor")
        return t

    def forward(self, r, diag, cE, cN, slot=2, z=None, add=False,
                features_unchanged=False):
        """z (+)= N(r) for fields of shape (B, nx, ny) (or (nx, ny)).  The
        operator features are recomputed unless ``features_unchanged`` (the
        stencil arrays are the ones of the previous call, unmodified)."""
        B, nx, ny = (1,) + tuple(r.shape) if r.dim() == 2 else tuple(r.shape)
        for t in (r, diag, cE, cN):
            self._f32(t)
        if not features_unchanged:
            L.check(L.lib().zj_mpnn_invalidate_features(self.handle),
                    "zj_mpnn_invalidate_features")
        out = torch.zeros_like(r) if z is None else self._f32(z)
        L.check(L.lib().zj_mpnn_forward(self.handle, B, nx, ny, _p(r),
                                        _p(diag), _p(cE), _p(cN), _p(out),
                                        int(add), int(slot), _stream()),
                "zj_mpnn_forward")
        return out

    def time_forward(self, r, diag, cE, cN, reps=20):
        """Average ms of one inference launch on an (nx, ny) grid (HIP
        events, back-to-back launches)."""
        nx, ny = r.shape
        z = torch.empty_like(r)
        ms = ctypes.c_float()
        L.check(L.lib().zj_mpnn_time_forward(
            self.handle, nx, ny, _p(self._f32(r)), _p(self._f32(diag)),
            _p(self._f32(cE)), _p(self._f32(cN)), _p(z), int(reps),
            ctypes.byref(ms), _stream()), "zj_mpnn_time_forward")
        return float(ms.value)

    def backward(self, dz, diag, slot=0, want_dr=True):
        B, nx, ny = (1,) + tuple(dz.shape) if dz.dim() == 2 else \
            tuple(dz.shape)
        dr = torch.empty_like(dz) if want_dr else None
        L.check(L.lib().zj_mpnn_backward(self.handle, B, nx, ny,
                                         _p(self._f32(dz)), _p(diag),
                                         _p(dr), int(slot), _stream()),
                "zj_mpnn_backward")
        return dr


# ───────────────────────── training data and training ─────────────────────
def _smooth_noise(n, gen, passes):
    x = torch.randn((1, 1, n, n), generator=gen, dtype=torch.float64,
                    device="cpu")
    k = torch.tensor([[0.0, 1.0, 0.0], [1.0, 4.0, 1.0], [0.0, 1.0, 0.0]],
                     dtype=torch.float64).reshape(1, 1, 3, 3) / 8.0
    for _ in range(passes):
        x = torch.nn.functional.conv2d(
            torch.nn.functional.pad(x, (1, 1, 1, 1), mode="replicate"), k)
    x = x.reshape(-1)
    return x / x.std()


def solver_pairs(n_pairs=5000, n=256, dx=0.0625, seeds=range(100, 110),
                 steps_per_seed=None, seed=0, log=None):
    """(b, h, rhs, p) pairs from the plain solver on n x n grids: the
    pressure equations of the stages of Chorin runs of the solitary-wave
    case (bed seeds ``seeds``, random wave positions), and for each of them
    one more right-hand side (white or smoothed noise) on the same operator.
    Every p solves S p = rhs to 1e-8 relative residual.  Device tensors."""
    from .fvm import ChorinSplitVAMSolver, ShallowWater
    from .fvm.chorin import fixed_dt
    from .mesh import Mesh
    seeds = list(seeds)
    if steps_per_seed is None:
        steps_per_seed = max(1, -(-n_pairs // (4 * len(seeds))))
    L_ = n * dx
    mesh = Mesh.quad(n, n, (0.0, L_, 0.0, L_)).set_bc("wall")
    model = ShallowWater(dimension=2, order=2, limiter="minmod",
                         riemann="hll")
    s = ChorinSplitVAMSolver(model, tol=1e-8, precond="mg")
    s.setup_simulation(mesh, n, n)
    gen = torch.Generator().manual_seed(seed)
    rng = np.random.default_rng(seed)
    dt = fixed_dt(dx, 1.0)
    pairs = []
    t0 = time.time()
    for sd in seeds:
        Q = torch.as_tensor(_wave_state(n, L_, sd, rng.uniform(0.15, 0.85)),
                            device=s.device)
        for _ in range(steps_per_seed):
            Q0 = Q
            for stage in range(2):
                Qs = s._hyperbolic_stage(dt, 0.0, Q)
                s.setup_operator(Qs[0], Qs[1])
                rhs = s.pressure_rhs(Qs, dt)
                p = s.solve(rhs)
                pairs.append((Qs[0].clone(), Qs[1].clone(), rhs, p))
                kind = len(pairs) % 4
                f = _smooth_noise(n, gen, 0 if kind < 2 else 4 + kind)
                f = f.to(s.device) * rhs.abs().max()
                pairs.append((Qs[0].clone(), Qs[1].clone(), f, s.solve(f)))
                Q = s._step_corrector_pure(Qs, p, dt,
                                           Q0 if stage == 1 else None)
                if len(pairs) >= n_pairs:
                    break
            if len(pairs) >= n_pairs:
                break
        if log:
            log(f"pairs {len(pairs)} ({time.time() - t0:.1f} s)")
        if len(pairs) >= n_pairs:
            break
    return pairs[:n_pairs], s


def error_pairs(s, pairs, n_pairs, seed=5):
    """(b, h, rhs, p) pairs whose right-hand sides are rhs = S e for random
    errors e (white, or smoothed by 1 or 3 passes) on the operators of
    ``pairs`` (every other one, in turn), p the plain solver's solution of
    S p = rhs (1e-8 relative residual; p equals e to that level): where the
    other solver pairs weight the smooth error the coarse correction
    already removes, these weight the oscillatory error the network acts
    on."""
    gen = torch.Generator().manual_seed(seed)
    ops = [(b, h) for (b, h, _, _) in pairs[::2]]
    n = int(round(ops[0][0].numel() ** 0.5))
    out = []
    for k in range(n_pairs):
        b, h = ops[k % len(ops)]
        s.setup_operator(b, h)
        e = _smooth_noise(n, gen, (0, 0, 1, 3)[k % 4]).to(s.device)
        rhs = s.apply(e)
        out.append((b, h, rhs, s.solve(rhs)))
    return out


def _wave_state(n, L_, seed, x0_frac):
    rng = np.random.default_rng(seed)
    kx = rng.integers(1, 5) / L_
    ky = rng.integers(1, 5) / L_
    ph = rng.uniform(0.0, 2.0 * np.pi)
    dx = L_ / n
    X, Y = np.meshgrid((np.arange(n) + 0.5) * dx, (np.arange(n) + 0.5) * dx,
                       indexing="ij")
    a, d = 0.2, 1.0
    k = np.sqrt(3.0 * a / (4.0 * d ** 3))
    c = np.sqrt(9.81 * (d + a))
    eta = a / np.cosh(k * (X - x0_frac * L_)) ** 2
    b = 0.1 * np.sin(2.0 * np.pi * (kx * X + ky * Y) + ph)
    Q = np.zeros((5, n * n))
    Q[0] = b.ravel()
    Q[1] = (d + eta - b).ravel()
    Q[2] = (c * eta).ravel()
    return Q


def vcycle_loss(s, net, b, h, rhs, p, train=True):
    """Energy-norm error of one learned V-cycle on S(h, b) p = rhs; with
    train, accumulates its parameter gradient into the network."""
    s.setup_operator(b, h)
    z = s.learned_vcycle(rhs.float(), train=train)
    e = z.double() - p
    Se = s.apply(e)
    E = float(torch.dot(p, rhs))
    loss = float(torch.dot(e, Se)) / E
    if train:
        s.learned_vcycle_backward((2.0 / E) * Se.float())
    return loss


def train(net: MPPreconditioner, pairs, s, epochs=8, batch=16, lr=3e-3,
          log=None, seed=0):
    """Adam on the V-cycle energy loss over the (b, h, rhs, p) pairs;
    returns the mean loss of every epoch."""
    s.set_preconditioner("learned", net)
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        tot = 0.0
        net.zero_grad()
        t0 = time.time()
        for k, i in enumerate(order):
            b, h, rhs, p = pairs[i]
            tot += vcycle_loss(s, net, b, h, rhs, p, train=True)
            if (k + 1) % batch == 0 or k + 1 == len(order):
                net.adam(lr)
                net.zero_grad()
        hist.append(tot / len(pairs))
        if log:
            log(f"epoch {ep} loss {hist[-1]:.5f} ({time.time() - t0:.1f} s)")
    return hist


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train the learned Chorin "
                                 "preconditioner on solver-generated pairs")
    ap.add_argument("--pairs", type=int, default=5000)
    ap.add_argument("--solver-pairs", type=int, default=3000,
                    help="of --pairs, how many come from Chorin runs and "
                         "noise right-hand sides; the rest from error "
                         "right-hand sides S e (all solved by the plain "
                         "solver)")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    pairs, s = solver_pairs(a.solver_pairs, a.n, log=log)
    pairs = pairs + error_pairs(s, pairs, a.pairs - len(pairs))
    net = MPPreconditioner(seed=0)
    hist = train(net, pairs, s, a.epochs, a.batch, a.lr, log=log)
    pathlib.Path(a.out).parent.mkdir(parents=True, exist_ok=True)
    np.savez(a.out, params=net.get_params(), hidden=HID, layers=NL,
             loss=np.array(hist), pairs=a.pairs, grid=a.n)
    log(f"saved {a.out}")


if __name__ == "__main__":
    main()
