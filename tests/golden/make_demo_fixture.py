"""Writes tests/golden/demo/demo.npz from the reference's demo data files.

Inputs (reference tree, data only): demo-data/velmod.csv (151 x 151 m/s),
demo-data/problemSetup.txt, demo-data/seismicShots/seis<ix>_151.csv
(9 shots, 800 x 151 float32), demo-result/final_image.csv and
demo-result/smoothed_velmod.csv (the outputs of run_demo.sh, `host` option).
The CSV values are stored unchanged (velocity float64, shots float32 as
main.py:122 reads them, results float64).

Run in a container that has the reference at /root/reference:
    python tests/golden/make_demo_fixture.py
"""
import glob
import os
import re

import numpy as np

REF = os.environ.get("WLSQ_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "demo", "demo.npz")


def main():
    vel = np.genfromtxt(os.path.join(REF, "demo-data", "velmod.csv"), delimiter=",")
    setup = open(os.path.join(REF, "demo-data", "problemSetup.txt")).read()
    files = glob.glob(os.path.join(REF, "demo-data", "seismicShots", "*.csv"))
    idx = [int(re.search(r"seis(.+?)_", os.path.basename(f)).group(1)) for f in files]
    order = np.argsort(idx)
    shots = np.stack([np.genfromtxt(files[i], delimiter=",", dtype=np.float32) for i in order])
    final = np.genfromtxt(os.path.join(REF, "demo-result", "final_image.csv"), delimiter=",")
    smooth = np.genfromtxt(os.path.join(REF, "demo-result", "smoothed_velmod.csv"), delimiter=",")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    np.savez_compressed(OUT, velocity=vel, setup=np.array(setup), shot_ix=np.array(idx)[order],
                        shots=shots, final_image=final, smoothed_velocity=smooth)
    print(OUT, os.path.getsize(OUT), "bytes", shots.shape)


if __name__ == "__main__":
    main()
