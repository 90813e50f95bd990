"""Writes kaufman_L16.json: the exact mean energy <E>(beta) of the 16x16 periodic Ising
ferromagnet (H = -sum_<ij> s_i s_j) at the 32 geometric betas 0.2-5.0 of the first benchmark
configuration (f32 betas, as PtIcmRunner::new casts them), from Kaufman's partition function
(tests/exact_ising.py).  Run from the repository root: python tests/golden/make_kaufman_L16.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from tests.exact_ising import mean_energy  # noqa: E402

# geometric_beta_schedule (tamc-core/src/sa.rs:13-22) on the f32 bounds, in f64, then cast to f32
lo, hi = np.log(np.float64(np.float32(0.2))), np.log(np.float64(np.float32(5.0)))
betas = [float(np.float32(np.exp(lo + k * (hi - lo) / 31))) for k in range(32)]
table = {"L": 16, "betas": betas, "mean_energy": [mean_energy(b, 16, 16) for b in betas]}
with open(os.path.join(HERE, "kaufman_L16.json"), "w") as f:
    json.dump(table, f, indent=1)
