"""Shared helpers for the parity tests."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    if den == 0:
        return num
    return num / den


def load_scene(sim_or_orc, s):
    sim_or_orc.set_particles(s["n_particles"], **s["particles"])
    if len(s["seeds"]):
        sim_or_orc.set_seed_cells(s["seeds"])


def say(*a):
    print(*a, flush=True)
