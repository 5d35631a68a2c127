"""Committed golden fixtures (small .npz files written by make_golden.py)."""
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def load_golden():
    cases = []
    for path in sorted(HERE.glob("case_*.npz")):
        with np.load(path, allow_pickle=False) as z:
            cases.append(dict(name=path.stem, N=[int(v) for v in z["N"]], P=[int(v) for v in z["P"]],
                              relaxed=bool(z["relaxed"]), x=z["x"], h=z["h"], y=z["y"]))
    return cases
