"""Writes the golden fixtures of tests/golden/ from the reference's own routines.

The routines are the reference's Fortran sources (src/oneline.f, cang.f,
asta.f, fillbessi0.f, besi1.f of spider-em/SPIDER) compiled by
``oracle/Makefile`` into ``oracle/_ref/libspider_ref.so``; this script only
calls them on seeded inputs and stores inputs and outputs.  Run from the repo
root (needs the reference tree at build time):

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libspider_ref.so not built (make -C oracle ref)")
    rng = np.random.default_rng(20261202)

    # CANG: doc-register angles -> DM(1..9)
    ang = rng.uniform(-360.0, 360.0, (256, 3)).astype(np.float32)
    ang[:8] = np.array([[0, 0, 0], [90, 0, 0], [0, 90, 0], [0, 0, 90], [30, 60, 45],
                        [180, 180, 180], [359.9, 0.1, 271.3], [-45, 135, 10]], np.float32)
    dm = np.stack([R.cang(a[0], a[1], a[2]) for a in ang])
    np.savez_compressed(os.path.join(OUT, "cang.npz"), phi_theta_psi=ang, dm=dm)

    # KB table (FILLBESSIL == the inline table of BP32F)
    tabs = {}
    for n2 in (128, 150, 256):
        tab, fltb, ln2, alpha, aaaa, nnn = R.fillbessil(n2)
        tabs[f"tab{n2}"] = tab
        tabs[f"meta{n2}"] = np.array([fltb, ln2, alpha, aaaa, nnn], np.float64)
    np.savez_compressed(os.path.join(OUT, "kbtab.npz"), **tabs)

    # Bessel functions used by the table and the window
    xs = np.concatenate([np.linspace(0.01, 6.5, 200), [5.85, 6.0, 6.25]]).astype(np.float32)
    np.savez_compressed(os.path.join(OUT, "bessel.npz"), x=xs,
                        besi1=np.array([R.besi1(float(x)) for x in xs], np.float32),
                        ribsl15=np.array([R.ribsl(float(x), 1.5) for x in xs], np.float32),
                        rjbsl15=np.array([R.rjbsl(float(x), 1.5) for x in xs], np.float32))

    # ASTA_D mean outside the circle, odd and even boxes
    imgs, means = [], []
    for L in (16, 17, 75):
        img = rng.standard_normal((L, L)).astype(np.float32) + 0.25
        aba, dklp = R.asta_d(img, float(L // 2))
        imgs.append(img)
        means.append([aba, dklp])
    np.savez_compressed(os.path.join(OUT, "asta.npz"), img16=imgs[0], img17=imgs[1], img75=imgs[2],
                        aba_dklp=np.array(means, np.float64))

    # ONELINE: all rows J = -NX+1..NX of one projection, serial order, on a
    # random BI and several orientations (including the edge-on and axis cases)
    cases = {}
    for ci, (nx, a) in enumerate([(12, (0.0, 0.0, 0.0)), (12, (17.0, 63.0, 211.0)),
                                  (12, (0.0, 90.0, 0.0)), (15, (123.4, 77.7, 5.5)),
                                  (16, (271.0, 145.0, 33.0)), (16, (45.0, 45.0, 45.0))]):
        n = 2 * nx
        tab, fltb = R.fillbessil(n)[:2]
        bi = (rng.standard_normal((n, nx + 1)) + 1j * rng.standard_normal((n, nx + 1))).astype(np.complex64)
        dmv = R.cang(*np.float32(a))
        X = np.zeros((n, n, nx + 1), np.complex64)
        W = np.zeros((n, n, nx + 1), np.float32)
        for J in range(-nx + 1, nx + 1):
            R.oneline(J, X, W, bi, dmv)
        cases[f"c{ci}_nx"] = np.array([nx])
        cases[f"c{ci}_dm"] = dmv
        cases[f"c{ci}_tab"] = tab
        cases[f"c{ci}_fltb"] = np.array([fltb], np.float32)
        cases[f"c{ci}_bi"] = bi
        cases[f"c{ci}_X"] = X
        cases[f"c{ci}_W"] = W
    np.savez_compressed(os.path.join(OUT, "oneline.npz"), **cases)

    # FSC shell sums (PR3DB) of two 24^3 volumes: transforms in SPIDER's LSD
    # layout (conjugated numpy rfftn, the SPIDER sign)
    n = 24
    v1 = rng.standard_normal((n, n, n)).astype(np.float32)
    v2 = (0.6 * v1 + 0.8 * rng.standard_normal((n, n, n))).astype(np.float32)
    fsc_out = {"v1": v1, "v2": v2}
    for tag, wi in (("w05", 0.5), ("w10", 1.0)):
        lsd = []
        for v in (v1, v2):
            F = np.conj(np.fft.rfftn(v.astype(np.float64)))
            a = np.zeros((n, n, n + 2), np.float32)
            a[:, :, 0::2] = F.real
            a[:, :, 1::2] = F.imag
            lsd.append(a)
        r = R.pr3db(lsd[0], lsd[1], n, wi=wi)
        for k in ("csum", "csum1", "csum2", "lr"):
            fsc_out[f"{tag}_{k}"] = r[k]
    np.savez_compressed(os.path.join(OUT, "fsc.npz"), **fsc_out)
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
