"""Writes the .in deck fixtures of tests/test_gpu_deck.py.

Two decks in the reference's format (keys of domain.cpp:133-196,
linearElastic.cpp:129-139, fault.cpp:63-101, strikeSlip_linearElastic_qd.cpp
:158-186) with their binary PETSc Vec inputs in the deck's inputDir
(big-endian classid 1211214, int32 length, float64 values;
genFuncs.cpp:74-94):

* deck_c1: the C1 recipe at 21x21 -- order 2, constant grid, mu = 36 GPa,
  aging law, a = 0.015 above 12 km ramping to 0.025 at 17 km, remote loading,
  guessSteadyStateICs = 1 with stateVals given so that the file psi (psi_ss
  at vL times 1 +- 1%, SplitMix64 seed 0x5C7C1E) is kept (fault.cpp:646-648);
* deck_c2: the C2 recipe at 25x25 -- order 4, variableGridSpacing, z from the
  file z (Lz sinh(5 r)/sinh(5)), bulk mu and cs from momBal_mu / momBal_cs
  (30 GPa at the surface smoothed to 36 GPa at 4 km), the rest as C1.

inputDir is relative to the repository root, as the reference resolves it
against the working directory.  Run from anywhere:
    python tests/golden/make_decks.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def splitmix64(seed, n):
    """n uniforms in [0, 1) from the 53 high bits of SplitMix64 outputs."""
    mask = (1 << 64) - 1
    s = seed & mask
    out = np.empty(n)
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def write_vec(path, v):
    v = np.ascontiguousarray(v, dtype=">f8")
    with open(path, "wb") as f:
        f.write((1211214).to_bytes(4, "big"))
        f.write(len(v).to_bytes(4, "big"))
        f.write(v.tobytes())


COMMON = """\
Ly = 30
Lz = 30
vL = 1e-9
bulkDeformationType = linearElastic
momentumBalanceType = quasidynamic
rhoVals = [3 3]
rhoDepths = [0 30]
stateLaw = agingLaw
f0 = 0.6
v0 = 1e-6
aVals = [0.015 0.015 0.025 0.025]
aDepths = [0 12 17 30]
bVals = [0.02 0.02]
bDepths = [0 30]
DcVals = [8e-3 8e-3]
DcDepths = [0 30]
sNVals = [50 50]
sNDepths = [0 30]
stateVals = [0.7 0.7]
stateDepths = [0 30]
guessSteadyStateICs = 1
momBal_bcR_qd = remoteLoading
momBal_bcT_qd = freeSurface
momBal_bcL_qd = symmFault
momBal_bcB_qd = freeSurface
timeIntegrator = RK32
timeControlType = PID
timeStepTol = 1e-8
timeIntInds = [psi slip]
normType = L2_absolute
initDeltaT = 1e-3
maxTime = 1e15
maxStepCount = 20
"""


def psi_file(Nz):
    f0, b, v0, vL = 0.6, 0.02, 1e-6, 1e-9
    psi_ss = f0 - b * np.log(vL / v0)
    return psi_ss * (1.0 + 0.02 * splitmix64(0x5C7C1E, Nz) - 0.01)


def main():
    d1 = os.path.join(HERE, "deck_c1")
    os.makedirs(d1, exist_ok=True)
    N = 21
    with open(os.path.join(d1, "c1.in"), "w") as f:
        f.write(f"Ny = {N}\nNz = {N}\norder = 2\ngridSpacingType = constantGridSpacing\n")
        f.write("inputDir = tests/golden/deck_c1/\n")
        f.write("muVals = [36 36]\nmuDepths = [0 30]\n")
        f.write(COMMON)
    write_vec(os.path.join(d1, "psi"), psi_file(N))

    d2 = os.path.join(HERE, "deck_c2")
    os.makedirs(d2, exist_ok=True)
    N = 25
    r = np.tile(np.linspace(0.0, 1.0, N), N)
    z = 30.0 * np.sinh(5.0 * r) / np.sinh(5.0)
    t = np.clip(z / 4.0, 0.0, 1.0)
    mu = np.where(z >= 4.0, 36.0, 30.0 + 6.0 * (3 * t ** 2 - 2 * t ** 3))
    with open(os.path.join(d2, "c2.in"), "w") as f:
        f.write(f"Ny = {N}\nNz = {N}\norder = 4\ngridSpacingType = variableGridSpacing\n")
        f.write("inputDir = tests/golden/deck_c2/\n")
        f.write("muVals = [36 36]\nmuDepths = [0 30]\n")
        f.write(COMMON)
    write_vec(os.path.join(d2, "psi"), psi_file(N))
    write_vec(os.path.join(d2, "z"), z)
    write_vec(os.path.join(d2, "momBal_mu"), mu)
    write_vec(os.path.join(d2, "momBal_cs"), np.sqrt(mu / 3.0))


if __name__ == "__main__":
    main()
