#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference's own data.

Run in a container where the reference checkout exists (default
/root/reference); the fixtures it writes are committed, so nothing at test or
bench time reads the reference.

Fixtures written (all plain JSON numbers, no code):
  init_gen_128.json   the inline vectors of test/test_init_gen.cpp:29-371
                      (inputs fc/psic/Up, the pack->unpack references and the
                      Yleft_ref / Yright_ref Taylor-series state vectors)
  lapack_10.json      the 10x10 system of test/test_lapack.cpp:29-46
  anchor_4D_<N>.json  the shipped near-solutions data/simulation_4D_<N>.json,
                      N = 128 ... 2048, reduced to the packed Newton vector x
                      (ICG.cpp:678-711 pack; NewtonSolver.cpp:273-274 puts
                      Delta in slot N/2+2) plus Delta and the solver keys.
"""
import argparse
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def cpp_arrays(text):
    """name -> list of numbers (or of (re, im) pairs) for every `vec_* name = {...};`."""
    out = {}
    for m in re.finditer(r"(vec_real|vec_complex|mat_real)\s+(\w+)\s*=\s*\{(.*?)\};", text, re.S):
        kind, name, body = m.groups()
        body = re.sub(r"//[^\n]*", "", body)
        if kind == "vec_complex":
            pairs = re.findall(r"\{\s*([-+0-9.eE]+)\s*,\s*([-+0-9.eE]+)\s*\}", body)
            out[name] = [[float(a), float(b)] for a, b in pairs]
        elif kind == "mat_real":
            rows = re.findall(r"\{([^{}]*)\}", body)
            out[name] = [[float(v) for v in r.split(",") if v.strip()] for r in rows]
        else:
            out[name] = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
    return out


def pack_numpy(odd1, odd2, even):
    """ICG.cpp:678-711 with numpy's FFT: forwardFFT is exp(+2 pi i jk/N)/N,
    which is numpy's ifft; halveModes keeps |k| < N/4 and folds +-N/4."""
    N = len(odd1)
    M, n = N // 2, 3 * N // 4

    def halved(x):
        X = np.fft.ifft(np.asarray(x, np.float64))
        h = np.zeros(M, np.complex128)
        h[: M // 2] = X[: M // 2]
        h[M // 2] = X[M // 2] + X[3 * M // 2]
        h[M // 2 + 1:] = X[M + M // 2 + 1:]
        return h

    o1, o2, ev = halved(odd1), halved(odd2), halved(even)
    Z = np.zeros(n)
    for j in range(n // 6):
        Z[2 * j], Z[2 * j + 1] = o1[2 * j + 1].real, o1[2 * j + 1].imag
        Z[2 * j + n // 3], Z[2 * j + 1 + n // 3] = o2[2 * j + 1].real, o2[2 * j + 1].imag
        Z[2 * j + 2 * n // 3], Z[2 * j + 1 + 2 * n // 3] = ev[2 * j].real, ev[2 * j].imag
    Z[2 * n // 3 + 1] = ev[n // 3].real
    return Z


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    ref = args.reference

    with open(os.path.join(ref, "test", "test_init_gen.cpp")) as f:
        arrs = cpp_arrays(f.read())
    fix = {
        "source": "test/test_init_gen.cpp:18-23 (parameters), 29-371 (vectors)",
        "Ntau": 128, "Dim": 4.0, "Delta": 3.445501184653556, "XLeft": 0.001, "XRight": 0.999,
    }
    for k in ("fc", "psic", "Up", "Up_unpack_ref", "psic_unpack_ref", "fc_unpack_ref",
              "Yleft_ref", "Yright_ref"):
        fix[k] = arrs[k]
    with open(os.path.join(HERE, "init_gen_128.json"), "w") as f:
        json.dump(fix, f)

    with open(os.path.join(ref, "test", "test_lapack.cpp")) as f:
        t = f.read()
    la = cpp_arrays(t)
    with open(os.path.join(HERE, "lapack_10.json"), "w") as f:
        json.dump({"source": "test/test_lapack.cpp:29-46", "A": la["mat"], "b": la["b"],
                   "x": la["ref"]}, f)

    for N in (128, 256, 512, 1024, 2048):
        with open(os.path.join(ref, "data", f"simulation_4D_{N}.json")) as f:
            d = json.load(f)
        ic = d["Initial_Conditions"]
        x = pack_numpy(ic["Up"], ic["psic"], ic["fc"])
        x[N // 2 + 2] = ic["Delta"]
        keys = {k: d[k] for k in ("Ntau", "Dim", "XLeft", "XMid", "XRight", "EpsNewton",
                                  "PrecisionNewton", "SlowError", "NLeft", "NRight",
                                  "SchemeIRK", "PrecisionIRK", "MaxIterIRK")}
        with open(os.path.join(HERE, f"anchor_4D_{N}.json"), "w") as f:
            json.dump({"source": f"data/simulation_4D_{N}.json", "config": keys,
                       "Delta": ic["Delta"], "x": x.tolist()}, f)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
