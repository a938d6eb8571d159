"""The smoothing step as a script with py/smbivspline.py's interface, for the
C++ driver's ScriptSmoother (the reference's _smoothPowerSpectra runs its
script the same way, core/quadratic_estimate.cpp:353-396):

    python3 -m lyspeq_amd.smooth InputPS OutputPS [--interp_log]

InputPS is the file writeSpectrumEstimates writes (core/quadratic_estimate.cpp:
624-650): a line "nz nk", a line of z-bin counts, then "z k P e" per band,
z-major; OutputPS gets one smoothed P per band and line, in the same order.
The fit is lyspeq_amd.driver.smooth_values."""
from __future__ import annotations

import argparse
import sys

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("InputPS")
    ap.add_argument("OutputPS")
    ap.add_argument("--interp_log", action="store_true")
    args = ap.parse_args(argv)
    from lyspeq_amd.driver import smooth_values
    z, k, p, e = np.genfromtxt(args.InputPS, delimiter=" ", skip_header=2, unpack=True)
    np.savetxt(args.OutputPS, smooth_values(z, k, p, e, args.interp_log), fmt="%.18e")
    return 0


if __name__ == "__main__":
    sys.exit(main())
