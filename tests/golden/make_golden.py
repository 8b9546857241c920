"""Writes the golden fixtures of the NCS cost, gradient and minimiser.

Independent of the oracle and of the product: the DFT comes from
numpy.fft (a library routine), the minimiser from scipy's L-BFGS-B run to a
tight tolerance, and the Parseval case from its closed form.  The cost
definition is the reference's (clib/pyCNCS/ncs.c:158-299, 335-364;
clib/pyCNCS/test/py_ref.py:12-58):

    f(u) = sum u - (d + gamma) ln max(u + gamma, 1e-6)
           + alpha / N^2 sum_kl M_kl^2 |fft2(u)_kl|^2          (M in FFT order)

Run:  python tests/golden/make_golden.py   (writes ncs_golden.npz here)
"""
import os

import numpy
import scipy.optimize

HERE = os.path.dirname(os.path.abspath(__file__))
N = 16


def otf_mask():
    """1 - circular-pupil OTF (NA 1.4, 700 nm, 100 nm pixels), centred."""
    X, Y = numpy.meshgrid(numpy.arange(-8, 8), numpy.arange(-8, 8))
    rho = numpy.sqrt(X ** 2 + Y ** 2) / (16 * 0.1) / (2 * 1.4 / 0.7)
    rc = numpy.minimum(rho, 1.0)
    T = (2 / numpy.pi) * (numpy.arccos(rc) - rc * numpy.sqrt(1 - rc ** 2))
    T[rho >= 1] = 0
    return 1 - T


def gaussian_mask():
    """opencl/pyOpenCLNCS/py_ref.py:129-137 (centred here)."""
    x = (numpy.arange(16) - 8.0) / 3.0
    gx = numpy.exp(-x * x)
    m = 1.0 - numpy.outer(gx, gx)
    m -= m.min()
    return m / m.max()


def random_mask(rng):
    """clib/pyCNCS/test/py_ref.py:77-85."""
    m = numpy.real(numpy.fft.fftshift(numpy.fft.fft2(rng.uniform(0, 1, (16, 16)))))
    m -= m.min()
    return m / m.max()


def cost_grad(u, d, g, mask_fft, alpha):
    u = u.reshape(N, N)
    v = numpy.maximum(u + g, 1e-6)
    ll = numpy.sum(u - (d + g) * numpy.log(v))
    U = numpy.fft.fft2(u)
    q = mask_fft ** 2
    nc = numpy.sum(q * numpy.abs(U) ** 2) / N ** 2
    gll = 1.0 - (d + g) / v
    gnc = 2.0 * numpy.real(numpy.fft.ifft2(q * U))
    return ll + alpha * nc, (gll + alpha * gnc).ravel()


def main():
    rng = numpy.random.default_rng(20190419)
    masks = {"otf": numpy.fft.fftshift(otf_mask()),
             "gauss": numpy.fft.fftshift(gaussian_mask()),
             "random": numpy.fft.fftshift(random_mask(rng))}
    out = {}
    for name, mf in masks.items():
        out["mask_" + name] = mf
    # 1. cost / gradient at random points, u in (-2, 10) exercises the guard
    ne = 12
    d = rng.uniform(0.01, 10.0, (ne, N, N))
    g = rng.uniform(2.0, 4.0, (ne, N, N))
    u = rng.uniform(-2.0, 10.0, (ne, N, N))
    out["eval_d"], out["eval_g"], out["eval_u"] = d, g, u
    for name, mf in masks.items():
        cs, gs = [], []
        for i in range(ne):
            c, gr = cost_grad(u[i], d[i], g[i], mf, 0.1)
            cs.append(c)
            gs.append(gr.reshape(N, N))
        out["eval_cost_" + name] = numpy.array(cs)
        out["eval_grad_" + name] = numpy.array(gs)
    # 2. minimisers (scipy L-BFGS-B with the analytic gradient, tight tolerance)
    ns = 6
    sd = rng.uniform(10.0, 20.0, (ns, N, N))
    sg = rng.uniform(2.0, 4.0, (ns, N, N))
    out["solve_d"], out["solve_g"] = sd, sg
    for name, mf in masks.items():
        for alpha in (0.02, 0.1):
            us, fs = [], []
            for i in range(ns):
                res = scipy.optimize.minimize(
                    cost_grad, sd[i].ravel(), args=(sd[i], sg[i], mf, alpha), jac=True,
                    method="L-BFGS-B",
                    options=dict(maxiter=5000, ftol=1e-15, gtol=1e-11, maxcor=20))
                us.append(res.x.reshape(N, N))
                fs.append(res.fun)
            key = "%s_a%g" % (name, alpha)
            out["solve_u_" + key] = numpy.array(us)
            out["solve_cost_" + key] = numpy.array(fs)
    # 3. Parseval: M = 1 makes the problem separable per pixel:
    #    1 - c/(u+g) + 2 alpha u = 0  ->  2 alpha u^2 + (1 + 2 alpha g) u + g - c = 0
    alpha = 0.1
    pd = rng.uniform(5.0, 60.0, (4, N, N))
    pg = rng.uniform(1.0, 20.0, (4, N, N))
    c = pd + pg
    out["parseval_d"], out["parseval_g"], out["parseval_alpha"] = pd, pg, alpha
    out["parseval_u"] = (-(1 + 2 * alpha * pg) + numpy.sqrt((1 + 2 * alpha * pg) ** 2
                                                           - 8 * alpha * (pg - c))) / (4 * alpha)
    numpy.savez_compressed(os.path.join(HERE, "ncs_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "ncs_golden.npz"))


if __name__ == "__main__":
    main()
