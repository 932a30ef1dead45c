"""Learned initial guess for the Stokes solve, end to end.

1. States: a variable-viscosity box (viscosity 10^(4 B), density rho0 - drho B
   from a seeded sum of Gaussian blobs B) and a "previous" state whose blobs
   sit slightly displaced, as between two time steps.
2. The plain method's own solves give the training targets: the previous and
   current Stokes solutions (host sparse direct solve of the same system the
   host Stokes solve iterates on) and the plain pressure guess of both states
   (initialPressureGuess on the GPU).
3. The U-Net (fvmic2d_amd.unet, hand-written MFMA kernels) maps
   [log10 eta / 4, (rho - rho0) / drho, vx_prev, vy_prev, q_prev] to the
   dynamic pressure q = p - p_plain on the node grid (the pressure minus the
   plain pressure guess); inputs and output are scaled by dataset constants.
   The guess is p = p_plain + q with the velocities that balance it: one
   solve with the velocity block the Stokes solve's preconditioner factors anyway.
4. The host Stokes solve: FGMRES + field split to a 1e-8 relative residual (host), from
   the plain guess as the reference uses it (zero velocity, pressure from
   initialPressureGuess), from the plain pressure with its balancing
   velocities, from the previous solution, and from the learned guess -- which
   falls back to the plain guess whenever its initial residual is larger.
"""
import time
from multiprocessing import get_context

import numpy as np
import scipy.sparse.linalg as sla

from . import stokes, workloads

G = 10.0
RHO0, DRHO = 3300.0, 100.0
CONTRAST = 1e4


class BoxGrid:
    def __init__(self, N, L=1.0):
        x = workloads.spacing_uniform(N, L)
        self.xp, self.xcp = workloads.pad_grid(x, L)
        self.yp, self.ycp = self.xp, self.xcp
        self.NX = self.NY = N
        self.LX = self.LY = L


def blob_field(N, seed, shift=0.0):
    """Normalised blob sum B in [0, 1] on the N x N nodes; shift displaces
    every blob by a seeded random vector of length `shift` (the previous state)."""
    r = np.random.default_rng(seed)
    nb = int(r.integers(4, 9))
    cx, cy = r.uniform(0.1, 0.9, nb), r.uniform(0.1, 0.9, nb)
    s = r.uniform(0.05, 0.2, nb)
    a = r.uniform(0.3, 1.0, nb)
    ang = r.uniform(0, 2 * np.pi, nb)
    x = np.linspace(0.0, 1.0, N)
    X, Y = np.meshgrid(x, x)
    B = np.zeros((N, N))
    for k in range(nb):
        ux = cx[k] + shift * np.cos(ang[k])
        uy = cy[k] + shift * np.sin(ang[k])
        B += a[k] * np.exp(-((X - ux) ** 2 + (Y - uy) ** 2) / s[k] ** 2)
    return B / B.max()


def fields(B):
    """etaS (nodes), etaN (cell centres, from the four surrounding nodes), mu, rho."""
    N = B.shape[0]
    etaS = CONTRAST ** B
    lg = np.log10(etaS)
    etaN = np.ones_like(etaS)
    etaN[1:, 1:] = 10.0 ** (0.25 * (lg[1:, 1:] + lg[:-1, 1:] + lg[1:, :-1] + lg[:-1, :-1]))
    mu = np.full((N, N), 1e99)  # viscous limit (Z = 1)
    rho = RHO0 - DRHO * B
    return etaS, etaN, mu, rho


def direct_solution(N, B):
    g = BoxGrid(N)
    etaS, etaN, mu, rho = fields(B)
    A, b, Kb, Kc = stokes.assemble(stokes.BoxGrid.of(g), etaS, etaN, mu, mu, rho, 1.0, G)
    return sla.spsolve(A.tocsc(), b), Kc


def _worker(args):
    N, seed, shift = args
    Bp = blob_field(N, seed, shift)
    B = blob_field(N, seed, 0.0)
    xp, _ = direct_solution(N, Bp)
    x, Kc = direct_solution(N, B)
    return seed, B, Bp, xp, x, Kc


def pressure_plain(ctx, grid, rho_dev, p_dev, Kc, N):
    """initialPressureGuess on the GPU, scaled to the Stokes unknown (p / Kcont)
    and re-pinned to 0 in cell (1, 1) like the Stokes solve's system."""
    from .host import initial_pressure_guess
    initial_pressure_guess(ctx, grid, rho_dev, G, rtol=1e-10, maxit=50000, p_out=p_dev)
    p = p_dev.get().reshape(N, N) / Kc
    return p - p[1, 1]


def enforce_box(x, N):
    """Give a guess the values its boundary rows prescribe: wall velocities 0,
    mirrored ghost rows, pressure ghost rows copied, pressure 0 in cell (1, 1)."""
    n = N * N
    vx = x[:n].reshape(N, N).copy()
    vy = x[n:2 * n].reshape(N, N).copy()
    p = x[2 * n:].reshape(N, N).copy()
    vx[:, 0] = 0.0
    vx[:, N - 1] = 0.0
    vx[N - 1, :] = vx[N - 2, :]
    vy[0, :] = 0.0
    vy[N - 1, :] = 0.0
    vy[:, N - 1] = vy[:, N - 2]
    p = p - p[1, 1]
    p[0, :] = p[1, :]
    p[:, 0] = p[:, 1]
    return np.concatenate([vx.ravel(), vy.ravel(), p.ravel()])


def generate(N, seeds, shift=0.03, workers=0, start="fork", progress=None):
    """The plain method's direct solves for every seed (host only).  With
    workers > 0 they run in worker processes: forked (start="fork", only before
    anything in this process has touched the GPU) or spawned (start="spawn",
    fresh interpreters that inherit no GPU state).  progress(done, total) is
    called after every 16 solves."""
    jobs = [(N, s, shift) for s in seeds]
    out = []
    if workers > 0:
        with get_context(start).Pool(workers) as pool:
            for r in pool.imap(_worker, jobs, chunksize=1):
                out.append(r)
                if progress and len(out) % 16 == 0:
                    progress(len(out), len(jobs))
        return out
    for j in jobs:
        out.append(_worker(j))
        if progress and len(out) % 16 == 0:
            progress(len(out), len(jobs))
    return out


class Dataset:
    """Samples at N x N nodes; the network sees the first H = N - 1 rows and columns."""

    def __init__(self, ctx, N, raw):
        from .lib import DeviceArray, DeviceGrid
        self.N, self.H = N, N - 1
        g = BoxGrid(N)
        dgrid = DeviceGrid(ctx, g)
        n = N * N
        rho_dev = DeviceArray(ctx, np.zeros(n))
        p_dev = DeviceArray(ctx, np.zeros(n))
        self.samples = []
        for seed, B, Bp, xp, x, Kc in raw:
            rho = fields(B)[3]
            rhop = fields(Bp)[3]
            rho_dev.set(rho)
            pp_cur = pressure_plain(ctx, dgrid, rho_dev, p_dev, Kc, N)
            rho_dev.set(rhop)
            pp_prev = pressure_plain(ctx, dgrid, rho_dev, p_dev, Kc, N)
            self.samples.append(dict(seed=seed, B=B, Bp=Bp, xprev=xp, x=x, Kc=Kc, pplain=pp_cur,
                                     pplain_prev=pp_prev))

    def arrays(self):
        """Raw (unscaled) inputs [S, H, H, 5] and targets q [S, H, H, 1]."""
        N, H, n = self.N, self.H, self.N * self.N
        S = len(self.samples)
        X = np.zeros((S, H, H, 5), np.float32)
        T = np.zeros((S, H, H, 1), np.float32)
        for k, s in enumerate(self.samples):
            X[k, ..., 0] = (np.log10(CONTRAST ** s["B"]) / 4.0)[:H, :H]
            X[k, ..., 1] = (-s["B"])[:H, :H]
            X[k, ..., 2] = s["xprev"][:n].reshape(N, N)[:H, :H]
            X[k, ..., 3] = s["xprev"][n:2 * n].reshape(N, N)[:H, :H]
            X[k, ..., 4] = (s["xprev"][2 * n:].reshape(N, N) - s["pplain_prev"])[:H, :H]
            T[k, ..., 0] = (s["x"][2 * n:].reshape(N, N) - s["pplain"])[:H, :H]
        return X, T


class LearnedGuess:
    """A trained U-Net plus the scales that map its channels to the Stokes unknowns.

    With residual=True (the default) the network learns the change of the
    dynamic pressure since the previous solution, q - q_prev, and the guess is
    q_prev plus its output: a network that has learned little then falls back
    towards the previous solution's pressure, not towards zero."""

    def __init__(self, ctx, levels=4, base=32, seed=0, residual=True):
        from .unet import UNet
        self.ctx = ctx
        self.net = UNet(ctx, levels=levels, base=base, cin=5, cout=1, seed=seed)
        self.vs = self.qs = self.dqs = 1.0
        self.residual = residual

    def _scaled(self, X, T=None):
        Xs = X.copy()
        Xs[..., 2:4] /= self.vs
        Xs[..., 4] /= self.qs
        if T is None:
            return Xs
        return Xs, T / self.dqs

    def train(self, X, T, epochs=30, batch=8, lr=2e-3, seed=0, log=None):
        from .unet import Buf, to_bf16_bits
        if self.residual:
            T = T - X[..., 4:5]
        # input scales from the previous solutions, output scale from the targets
        self.vs = float(np.sqrt(np.mean(X[..., 2:4] ** 2))) or 1.0
        self.qs = float(np.sqrt(np.mean(X[..., 4] ** 2))) or 1.0
        self.dqs = float(np.sqrt(np.mean(T ** 2))) or 1.0
        Xs, Ts = self._scaled(X, T)
        S, H, W, _ = Xs.shape
        cin = self.net.cin
        xb = np.zeros((batch, H, W, cin), np.float32)
        dX = Buf(self.ctx, (batch, H, W, cin), np.uint16)
        dT = Buf(self.ctx, (batch, H, W, 1), np.float32)
        rng = np.random.default_rng(seed)
        steps = max(1, S // batch)
        total = epochs * steps
        hist = []
        t0 = time.time()
        it = 0
        for ep in range(epochs):
            order = rng.permutation(S)
            el = 0.0
            for s in range(steps):
                idx = order[s * batch:(s + 1) * batch]
                xb[...] = 0.0
                xb[..., :5] = Xs[idx]
                dX.set(to_bf16_bits(xb))
                dT.set(Ts[idx])
                frac = it / total
                self.net.forward(dX)
                el += self.net.loss(dT)
                self.net.backward()
                self.net.adam(lr=lr * 0.5 * (1.0 + np.cos(np.pi * frac)))
                it += 1
            hist.append(el / steps)
            if log:
                log("epoch %d loss %.4e (%.1f s)" % (ep, hist[-1], time.time() - t0))
        self.train_seconds = time.time() - t0
        return hist

    def predict(self, X):
        """Raw inputs [S, H, W, 5] -> predicted dynamic pressure q [S, H, W, 1]."""
        from .unet import Buf, to_bf16_bits
        S, H, W, _ = X.shape
        Xs = self._scaled(X)
        out = np.zeros((S, H, W, 1), np.float32)
        dX = Buf(self.ctx, (1, H, W, self.net.cin), np.uint16)
        xb = np.zeros((1, H, W, self.net.cin), np.float32)
        for k in range(S):
            xb[...] = 0.0
            xb[0, ..., :5] = Xs[k]
            dX.set(to_bf16_bits(xb))
            out[k] = self.net.forward(dX).get()[0, ..., :1]
        out = out * self.dqs
        if self.residual:
            out = out + X[..., 4:5]
        return out


def pressure_from_output(o, sample, N):
    """p_plain + predicted q on the first N-1 rows / columns (the last row and
    column repeat their neighbours)."""
    H = N - 1
    q = np.zeros((N, N))
    q[:H, :H] = o[..., 0]
    q[H, :H] = q[H - 1, :H]
    q[:, H] = q[:, H - 1]
    return sample["pplain"] + q


def balanced(A, b, M, p, N):
    """(v, p) with v solving the momentum rows for the given pressure
    (A_uu v = b_u - A_up p, through the preconditioner's LU), boundary rows
    enforced."""
    n = N * N
    x = enforce_box(np.concatenate([np.zeros(2 * n), p.ravel()]), N)
    v = M.lu.solve(b[:2 * n] - M.Aup @ x[2 * n:])
    return enforce_box(np.concatenate([v, x[2 * n:]]), N)


def evaluate(ds, model, rtol=1e-8, restart=100, progress=None):
    """FGMRES iterations from the plain guess, the previous solution and the
    learned guess (with fallback) for every sample of ds; progress(k, total)
    after every sample."""
    N, n = ds.N, ds.N * ds.N
    X, _ = ds.arrays()
    pred = model.predict(X)
    g = stokes.BoxGrid.of(BoxGrid(N))
    rows = []
    for k, s in enumerate(ds.samples):
        etaS, etaN, mu, rho = fields(s["B"])
        A, b, Kb, Kc = stokes.assemble(g, etaS, etaN, mu, mu, rho, 1.0, G)
        zN = stokes.viscosity_z(etaS, etaN, mu, mu, 1.0)[1]
        M = stokes.FieldSplit(A, Kc, zN)
        plain = enforce_box(np.concatenate([np.zeros(2 * n), s["pplain"].ravel()]), N)
        plain_bal = balanced(A, b, M, s["pplain"], N)
        prev = s["xprev"]
        learned = balanced(A, b, M, pressure_from_output(pred[k], s, N), N)
        r0 = {nm: np.linalg.norm(b - A @ v) / np.linalg.norm(b) for nm, v in
              (("plain", plain), ("plain_balanced", plain_bal), ("prev", prev), ("learned", learned))}
        fallback = r0["learned"] > r0["plain"]
        res = {}
        for nm, v in (("plain", plain), ("plain_balanced", plain_bal), ("prev", prev),
                      ("learned", plain if fallback else learned), ("learned_no_fallback", learned)):
            _, its, rel = stokes.fgmres(A, b, v, M, rtol=rtol, restart=restart, maxit=3000)
            res[nm] = its
        rows.append(dict(seed=s["seed"], r0=r0, its=res, fallback=fallback))
        if progress:
            progress(k + 1, len(ds.samples))
    return rows


KINDS = ("plain", "plain_balanced", "prev", "learned")


def summarize(rows):
    its = {k: float(np.mean([r["its"][k] for r in rows])) for k in KINDS + ("learned_no_fallback",)}
    r0 = {k: float(np.exp(np.mean([np.log(r["r0"][k]) for r in rows]))) for k in KINDS}
    return {"mean_iterations": its, "geomean_initial_relres": r0,
            "reduction_vs_plain": 1.0 - its["learned"] / its["plain"],
            "reduction_vs_plain_balanced": 1.0 - its["learned"] / its["plain_balanced"],
            "reduction_vs_previous": 1.0 - its["learned"] / its["prev"],
            "fallbacks": int(sum(r["fallback"] for r in rows)), "samples": len(rows)}
