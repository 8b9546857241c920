"""Learned linear start for the NCS solver, in the real-DFT basis.

The solver's iterations converge linearly from the plain start u0 = d; a
better start saves whole iterations.  This predictor is linear in two
normalised input channels and diagonal in the orthonormal real-DFT basis R
of the 16 x 16 tile (rows cos k, k = 0..8, then sin k, k = 1..7; Z = R X R^T),
the basis the solver already works in:

    mu = mean d,  sigma = sqrt(max(mu, 1) + mean gamma),
    x0 = (d - mu) / sigma,  x1 = gamma / sigma^2,
    Z(u0) = Z(d) + sigma (A * Z(x0) + B * Z(x1) + c Z(1))      (coefficient-wise)

Because R is orthonormal, the least-squares fit of (u* - d) / sigma over
training tiles separates into one 3-parameter regression per coefficient,
solved in closed form.  The training pairs (tile, u*) come from the plain
GPU solver on synthetic frames with seeds disjoint from the benchmark's.
The solve kernel evaluates the predictor (NCS_START_LINEAR) with three basis
changes and keeps the plain start per tile when it scores the lower cost.
"""
import argparse
import json
import os
import sys
import time

import numpy

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_COEF = os.path.join(HERE, "weights", "linstart_otf_a0.1.npy")
DEFAULT_SLOPE = os.path.join(HERE, "weights", "linstart_otf_slope.npy")
ALPHA_REF = 0.1  # the alpha DEFAULT_COEF was fitted at
SLOPE_SPAN = numpy.log(8.0)  # slopes extrapolate at most 8x from alpha_ref
N_COEF = 2 * 256 + 1


def real_dft_basis(n=16):
    """Orthonormal real DFT basis R [n, n]: rows cos k (k = 0..n/2), then
    sin k (k = 1..n/2-1); the same basis the solver's tables hold."""
    i = numpy.arange(n)
    R = numpy.zeros((n, n))
    h = n // 2
    for a in range(n):
        if a == 0:
            R[a] = 1.0 / numpy.sqrt(n)
        elif a == h:
            R[a] = numpy.where(i % 2 == 1, -1.0, 1.0) / numpy.sqrt(n)
        elif a < h:
            R[a] = numpy.sqrt(2.0 / n) * numpy.cos(2.0 * numpy.pi * a * i / n)
        else:
            R[a] = numpy.sqrt(2.0 / n) * numpy.sin(2.0 * numpy.pi * (a - h) * i / n)
    return R


_R = real_dft_basis()


def spec(x):
    """Z = R X R^T of tiles [..., 16, 16]."""
    return _R @ x @ _R.T


def pix(z):
    """X = R^T Z R, the inverse of spec."""
    return _R.T @ z @ _R


def features(d, g):
    """(Z(x0), Z(x1), sigma) of tiles [n, 16, 16] (float64)."""
    d = numpy.asarray(d, numpy.float64)
    g = numpy.asarray(g, numpy.float64)
    mu = d.mean(axis=(1, 2), keepdims=True)
    sig = numpy.sqrt(numpy.maximum(mu, 1.0) + g.mean(axis=(1, 2), keepdims=True))
    return spec((d - mu) / sig), spec(g / sig ** 2), sig


def fit(d, g, u, ridge=1e-9):
    """Closed-form least squares of Z((u - d) / sigma) on (Z(x0), Z(x1), Z(1)),
    one 3 x 3 system per coefficient.  Returns the N_COEF coefficients
    [A (16x16), B (16x16), c] as float32."""
    X0, X1, sig = features(d, g)
    Y = spec((numpy.asarray(u, numpy.float64) - numpy.asarray(d, numpy.float64)) / sig)
    one = spec(numpy.ones((16, 16)))
    A = numpy.zeros((16, 16))
    B = numpy.zeros((16, 16))
    c = 0.0
    for a in range(16):
        for b in range(16):
            cols = [X0[:, a, b], X1[:, a, b]]
            if abs(one[a, b]) > 0.5:
                cols.append(numpy.full(len(X0), one[a, b]))
            M = numpy.stack(cols, 1)
            G = M.T @ M + ridge * numpy.eye(M.shape[1]) * len(M)
            w = numpy.linalg.solve(G, M.T @ Y[:, a, b])
            A[a, b], B[a, b] = w[0], w[1]
            if len(w) == 3:
                c = w[2]
    return numpy.concatenate([A.ravel(), B.ravel(), [c]]).astype(numpy.float32)


def at_alpha(coef, slope, alpha, alpha_ref=ALPHA_REF):
    """The coefficients a frame at alpha uses, as the solver forms them:
    coef + ell * slope with ell = ln(alpha / alpha_ref) clamped to
    +-SLOPE_SPAN; coef itself without slopes, at alpha_ref or for alpha <= 0."""
    coef = numpy.asarray(coef, numpy.float64)
    if slope is None or alpha == alpha_ref or not alpha > 0:
        return coef
    ell = float(numpy.clip(numpy.log(alpha / alpha_ref), -SLOPE_SPAN, SLOPE_SPAN))
    return coef + ell * numpy.asarray(slope, numpy.float64)


def predict(d, g, coef, alpha=None, slope=None, alpha_ref=ALPHA_REF):
    """The start u0 of tiles [n, 16, 16] (float64), as the solver computes it;
    with slope and alpha, at that alpha's coefficients."""
    if slope is not None and alpha is not None:
        coef = at_alpha(coef, slope, alpha, alpha_ref)
    coef = numpy.asarray(coef, numpy.float64)
    A = coef[:256].reshape(16, 16)
    B = coef[256:512].reshape(16, 16)
    c = coef[512]
    X0, X1, sig = features(d, g)
    one = spec(numpy.ones((16, 16)))
    return numpy.asarray(d, numpy.float64) + sig * pix(A * X0 + B * X1 + c * one)


def fit_slope(d, g, solutions, coef, alpha_ref=ALPHA_REF, ridge=1e-9):
    """Slopes in ln(alpha) of the coefficients, given the alpha_ref fit: for
    each coefficient, the least squares of the alpha_ref start's residual
    Z((u*_alpha - d) / sigma) - (A Z(x0) + B Z(x1) + c Z(1)) on
    ln(alpha / alpha_ref) [Z(x0), Z(x1), Z(1)] over every alpha of
    solutions, a list of (alpha, u* [n, 16, 16]) for the same tiles.
    Returns N_COEF float32 slopes [A', B', c'] in coef's layout."""
    coef = numpy.asarray(coef, numpy.float64)
    A0 = coef[:256].reshape(16, 16)
    B0 = coef[256:512].reshape(16, 16)
    c0 = coef[512]
    X0, X1, sig = features(d, g)
    one = spec(numpy.ones((16, 16)))
    d64 = numpy.asarray(d, numpy.float64)
    rows = []
    for alpha, u in solutions:
        Y = spec((numpy.asarray(u, numpy.float64) - d64) / sig)
        rows.append((numpy.log(alpha / alpha_ref), Y - (A0 * X0 + B0 * X1 + c0 * one)))
    A1 = numpy.zeros((16, 16))
    B1 = numpy.zeros((16, 16))
    c1 = 0.0
    for a in range(16):
        for b in range(16):
            cols, ys = [], []
            for ell, R in rows:
                c = [ell * X0[:, a, b], ell * X1[:, a, b]]
                if abs(one[a, b]) > 0.5:
                    c.append(numpy.full(len(X0), ell * one[a, b]))
                cols.append(numpy.stack(c, 1))
                ys.append(R[:, a, b])
            M = numpy.concatenate(cols)
            G = M.T @ M + ridge * numpy.eye(M.shape[1]) * len(M)
            w = numpy.linalg.solve(G, M.T @ numpy.concatenate(ys))
            A1[a, b], B1[a, b] = w[0], w[1]
            if len(w) == 3:
                c1 = w[2]
    return numpy.concatenate([A1.ravel(), B1.ravel(), [c1]]).astype(numpy.float32)


def load_slope(path=DEFAULT_SLOPE):
    """The committed slopes, or None when the file is absent."""
    if not os.path.exists(path):
        return None
    return load(path)


def save(coef, path=DEFAULT_COEF):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    numpy.save(path, numpy.asarray(coef, numpy.float32))


def load(path=DEFAULT_COEF):
    c = numpy.load(path, allow_pickle=False)
    if c.shape != (N_COEF,):
        raise ValueError("expected %d coefficients, found %s" % (N_COEF, c.shape))
    return c


def set_context_coef(coef, device=0, slope=None, alpha_ref=ALPHA_REF):
    """Hand the coefficients (and their alpha slopes) to the library
    (NCS_START_LINEAR)."""
    _lib.context(device).set_linear_start(coef, slope, alpha_ref)


def main(argv=None):
    ap = argparse.ArgumentParser(description="fit the learned linear start")
    sub = ap.add_subparsers(dest="cmd", required=True)
    f = sub.add_parser("fit")
    f.add_argument("--tiles", type=int, default=200_000)
    f.add_argument("--alpha", type=float, default=0.1)
    f.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    f.add_argument("--out", default=DEFAULT_COEF)
    fs = sub.add_parser("fit-slope", help="slopes in ln(alpha) for a fitted set")
    fs.add_argument("--tiles", type=int, default=60_000)
    fs.add_argument("--alphas", default="0.05,0.2,0.3")
    fs.add_argument("--coef", default=DEFAULT_COEF)
    fs.add_argument("--alpha-ref", type=float, default=ALPHA_REF)
    fs.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    fs.add_argument("--out", default=DEFAULT_SLOPE)
    args = ap.parse_args(argv)
    if args.cmd == "fit-slope":
        return _fit_slope_main(args)
    from . import warmstart
    log = lambda *a: print(*a, flush=True)
    d, g = warmstart.gen_windows(args.tiles, workers=args.workers, log=log)
    de, ge = warmstart.gen_windows(4096, seed0=20_000_000, workers=args.workers, log=log)
    t0 = time.time()
    td, tg, tu = warmstart.solve_pairs(d, g, alpha=args.alpha, log=log)
    u = tu.cpu().numpy()
    coef = fit(d, g, u)
    log("fit: %d tiles in %.1f s" % (len(d), time.time() - t0))
    save(coef, args.out)
    # held-out iterations, plain vs linear start, on the GPU solver
    ctx = _lib.context(0)
    ctx.set_linear_start(coef)
    pp = _lib.default_params()
    pl = _lib.default_params(start=_lib.START_LINEAR)
    _, it_p, st_p, c_p = ctx.solve_tiles(de, ge, args.alpha, pp)
    _, it_l, st_l, c_l = ctx.solve_tiles(de, ge, args.alpha, pl)
    rel = numpy.abs(c_l.astype(numpy.float64) - c_p) / numpy.abs(c_p)
    print(json.dumps({"coef": args.out, "tiles": len(d), "alpha": args.alpha,
                      "held_out_iterations_plain": float(it_p.mean()),
                      "held_out_iterations_linear": float(it_l.mean()),
                      "held_out_cost_max_rel": float(rel.max()),
                      "status_linear": {str(k): int(v) for k, v in
                                        zip(*numpy.unique(st_l, return_counts=True))}}))
    return 0


def _fit_slope_main(args):
    """Solutions of the same training windows at each alpha on the GPU, the
    slopes fitted to them, and held-out iterations at each alpha (history 1,
    the default initial Hessian) with the alpha_ref coefficients alone and
    with the slopes."""
    from . import warmstart
    log = lambda *a: print(*a, flush=True)
    alphas = [float(a) for a in args.alphas.split(",")]
    coef = load(args.coef)
    d, g = warmstart.gen_windows(args.tiles, workers=args.workers, log=log)
    de, ge = warmstart.gen_windows(4096, seed0=20_000_000, workers=args.workers, log=log)
    t0 = time.time()
    sols = []
    for a in alphas:
        _, _, tu = warmstart.solve_pairs(d, g, alpha=a, log=log)
        sols.append((a, tu.cpu().numpy()))
    slope = fit_slope(d, g, sols, coef, args.alpha_ref)
    log("fit-slope: %d tiles x %d alphas in %.1f s" % (len(d), len(alphas), time.time() - t0))
    save(slope, args.out)
    ctx = _lib.context(0)
    report = {"slope": args.out, "tiles": len(d), "alphas": alphas, "alpha_ref": args.alpha_ref}
    for a in alphas + [args.alpha_ref]:
        row = {}
        for label, sl in (("fixed", None), ("adapted", slope)):
            ctx.set_linear_start(coef, sl, args.alpha_ref)
            _, it, st, c = ctx.solve_tiles(de, ge, a, _lib.default_params(
                start=_lib.START_LINEAR, history=1))
            row[label] = {"iterations": float(it.mean()),
                          "status": {str(k): int(v) for k, v in
                                     zip(*numpy.unique(st, return_counts=True))}}
        report["held_out_alpha_%g" % a] = row
    print(json.dumps(report))
    return 0


if __name__ == "__main__":
    sys.exit(main())
