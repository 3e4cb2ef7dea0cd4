#!/usr/bin/env python3
"""Write the golden fixtures the CPU tests check the oracle against.

Two files, both plain JSON data (inputs and expected outputs):

reference_tables.json
    Numbers the reference itself holds, lifted as data from its source text
    (nothing of the reference is imported or run):
      - LamInt[256], exact Lambda(k) for k < 256      src/C/lambda.c:102-188
      - BMe[171] / BMeT[171], exact B(0) at M = 18    src/C/leg2cheb.c:6-123
      - the N = 20 all-ones known answers, L2C and C2L  src/bin/l2c.c:69-74
      - the Lambda known answers Lambda(0..2), Lambda(24.5)  src/bin/l2c.c:695-707
    Needs the reference checkout (default /root/reference, or --reference).

mp_fixtures.json
    Independent 40-digit mpmath values that pin what the reference's tables
    leave open:
      - Lambda(z) = Gamma(z+1/2)/Gamma(z+1) at points on both sides of every
        series cut-off and in the upward-recursion range;
      - L2C of seeded vectors, computed from the definition (evaluate the
        Legendre series at Chebyshev-Gauss points with mpmath's legendre(),
        then the exact discrete Chebyshev transform), and C2L as the exact
        triangular inverse of that matrix.  Neither uses the Lambda-product
        closed form the oracle implements, so a dropped term, a wrong sign or
        an index slip in the oracle shows up against them.

    python tests/golden/make_golden.py [--reference DIR]
"""
import argparse
import json
import os
import re

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def c_array(text, name):
    """Values of the C array initialiser `name[...] = { ... };` in `text`."""
    m = re.search(r"\b" + re.escape(name) + r"\s*\[[^\]]*\][^=]*=\s*\{(.*?)\}\s*;", text, re.S)
    if not m:
        raise ValueError(f"array {name} not found")
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]


def reference_tables(ref):
    lam = open(os.path.join(ref, "src/C/lambda.c")).read()
    l2c = open(os.path.join(ref, "src/C/leg2cheb.c")).read()
    cli = open(os.path.join(ref, "src/bin/l2c.c")).read()
    acc = cli[cli.index("void test_accuracy"):]
    lt = cli[cli.index("void test_lambda"):]
    exact245 = re.search(r"Lambda\(24\.5\) exact = %2\.18e\\n\",\s*([0-9.eE+-]+)\)", lt)
    out = {
        "LamInt": {"source": "src/C/lambda.c:102-188", "values": c_array(lam, "LamInt")},
        "BMe": {"source": "src/C/leg2cheb.c:6-63", "layout": "B(0) lower triangle, row-major",
                "values": c_array(l2c, "BMe")},
        "BMeT": {"source": "src/C/leg2cheb.c:65-123", "layout": "B(0)^T upper triangle, row-major",
                 "values": c_array(l2c, "BMeT")},
        "kat_n20_ones": {"source": "src/bin/l2c.c:69-74,94,104", "N": 20, "input": "all ones",
                         "l2c_first5": c_array(acc, "a"), "c2l_first5": c_array(acc, "b"),
                         "tol_l2": 1e-7},
        "lambda_kat": {"source": "src/bin/l2c.c:695-707",
                       "z": [0.0, 1.0, 2.0], "values": c_array(lt, "z0"), "tol": 1e-15,
                       "z245_exact": float(exact245.group(1)) if exact245 else None},
    }
    assert len(out["LamInt"]["values"]) == 256
    assert len(out["BMe"]["values"]) == 171 and len(out["BMeT"]["values"]) == 171
    assert len(out["kat_n20_ones"]["l2c_first5"]) == 5
    return out


def splitmix_uniform(n, seed):
    """U(-1,1) doubles from a SplitMix64 counter stream (same generator as the
    package's splitmix_uniform)."""
    mask = (1 << 64) - 1
    x = seed & mask
    out = np.empty(n)
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        out[i] = 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0
    return out


def l2c_matrix_mp(N):
    """L[n][j]: Chebyshev coefficient n of P_j, from N-point Chebyshev-Gauss
    collocation (exact for degree < N)."""
    xs = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / N) for k in range(N)]
    P = [[mp.legendre(j, x) for j in range(N)] for x in xs]
    L = mp.matrix(N, N)
    for n in range(N):
        Tn = [mp.cos(n * mp.pi * (k + mp.mpf(1) / 2) / N) for k in range(N)]
        c = mp.mpf(1) / N if n == 0 else mp.mpf(2) / N
        for j in range(n, N):
            L[n, j] = c * mp.fsum(Tn[k] * P[k][j] for k in range(N))
    return L


def mp_fixtures():
    mp.mp.dps = 40
    zs = [0.0, 0.25, 0.5, 1.0, 2.0, 3.5, 7.75, 9.5, 9.999, 10.0, 10.5, 13.9, 14.0, 14.5,
          19.5, 20.0, 20.5, 24.5, 39.5, 40.0, 40.5, 100.25, 140.5, 141.0, 141.5, 1000.5,
          2013.5, 2014.0, 2014.5, 3000.0, 65536.5, 1.0e6 + 0.5, 1.0e9]
    lam = [mp.nstr(mp.gamma(mp.mpf(z) + mp.mpf(1) / 2) / mp.gamma(mp.mpf(z) + 1), 30) for z in zs]
    cases = []
    for N, seed, decay in ((64, 7, 0.0), (130, 11, 1.0), (200, 3, 0.0)):
        u = splitmix_uniform(N, seed) / (np.arange(N) + 1.0) ** decay
        L = l2c_matrix_mp(N)
        um = mp.matrix([mp.mpf(float(v)) for v in u])
        b = L * um
        # C2L of the same input: solve L x = u (upper triangular, exact inverse)
        x = mp.matrix(N, 1)
        for n in range(N - 1, -1, -1):
            acc = um[n] - mp.fsum(L[n, j] * x[j] for j in range(n + 1, N))
            x[n] = acc / L[n, n]
        cases.append({"N": N, "seed": seed, "decay": decay, "input": [float(v) for v in u],
                      "l2c": [mp.nstr(b[i], 30) for i in range(N)],
                      "c2l": [mp.nstr(x[i], 30) for i in range(N)]})
    return {"generator": "mpmath %s, %d digits" % (mp.__version__, mp.mp.dps),
            "lambda": {"z": zs, "values": lam}, "transforms": cases}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--skip-mp", action="store_true")
    args = ap.parse_args()
    if os.path.isdir(args.reference):
        with open(os.path.join(HERE, "reference_tables.json"), "w") as f:
            json.dump(reference_tables(args.reference), f, indent=1)
    else:
        print(f"{args.reference} not found: reference_tables.json left as it is")
    if not args.skip_mp:
        with open(os.path.join(HERE, "mp_fixtures.json"), "w") as f:
            json.dump(mp_fixtures(), f, indent=1)


if __name__ == "__main__":
    main()
