"""Dynamic-aperture parameter scans: the reference's HTCondor driver
(htcondor_jobs/compose_config.py, htcondor_jobs/run_sim.py) on one GPU.

The reference writes one line per configuration to configs/all_jobs.txt and runs
one job per line: a num_points x num_points grid over [0, max_radius]^2 (px = py
= 0), an "sps" tracker with the line's tunes and epsilon, track(num_iterations,
mu, barrier=10), and the survival turns saved as dataset "steps" of
stability_henon_{omega_x}_{omega_y}_{epsilon}_{mu}_{max_radius}.h5 (decimal
points written as "d"). Here the same lines run as batched jobs
(henon_map_cpp_amd.track_batch), as many per launch sequence as their
modulation tables fit in a memory budget, and each job's steps are saved as the
reference saves them: dataset "steps" (uint32, little-endian, contiguous) of
the .h5 file of that name, written through the HDF5 C library
(henon_map_cpp_amd.h5; fmt="npy" writes numpy's format instead).

Command line, as run_sim.py (one configuration) or for a whole job list:
    python -m henon_map_cpp_amd.scan --omegax 0.168 --omegay 0.201 --epsilon 32 --mu 0.5
    python -m henon_map_cpp_amd.scan --jobs configs/all_jobs.txt --num_iterations 100000
"""
from __future__ import annotations

import argparse
import datetime
import itertools
import os
from dataclasses import dataclass

import numpy as np

# htcondor_jobs/compose_config.py:3-9
OMEGA_LIST = [(0.168, 0.201, 0.75), (0.310, 0.320, 0.25)]
EPSILON_LIST = [0.0, 1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0, 128.0, 256.0]
MU_LIST = [0.0, 0.25, 0.5, 0.75, 1.0]


def f_to_str(f):
    """htcondor_jobs/run_sim.py:10-12: the float's str with "." -> "d"."""
    return str(f).replace(".", "d")


@dataclass(frozen=True)
class ScanConfig:
    omega_x: float
    omega_y: float
    epsilon: float
    mu: float
    max_radius: float

    def file_stem(self):
        """htcondor_jobs/run_sim.py:58-60."""
        return (f"stability_henon_{f_to_str(self.omega_x)}_{f_to_str(self.omega_y)}_{f_to_str(self.epsilon)}_"
                f"{f_to_str(self.mu)}_{f_to_str(self.max_radius)}")


def compose_config(path="configs/all_jobs.txt", omega_list=OMEGA_LIST, epsilon_list=EPSILON_LIST, mu_list=MU_LIST):
    """htcondor_jobs/compose_config.py:11-18: every (omega, epsilon, mu)
    combination, one "omega_x omega_y epsilon mu max_radius" line each."""
    param_list = list(itertools.product(omega_list, epsilon_list, mu_list))
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        for (omega_x, omega_y, max_radius), epsilon, mu in param_list:
            f.write(f"{omega_x} {omega_y} {epsilon} {mu} {max_radius}\n")
    return [ScanConfig(ox, oy, e, m, r) for (ox, oy, r), e, m in param_list]


def read_config(path):
    """The job list written by compose_config."""
    out = []
    with open(path) as f:
        for line in f:
            if line.strip():
                ox, oy, e, m, r = (float(v) for v in line.split())
                out.append(ScanConfig(ox, oy, e, m, r))
    return out


def initial_conditions(max_radius, num_points):
    """htcondor_jobs/run_sim.py:27-35."""
    positions = np.linspace(0, max_radius, num_points)
    xx, yy = np.meshgrid(positions, positions)
    x = xx.flatten()
    y = yy.flatten()
    return x, np.zeros_like(x), y, np.zeros_like(y)


def save_steps(path_stem, steps, fmt="h5"):
    """run_sim.py:55-59: `steps` as dataset "steps" of path_stem + ".h5" (or
    numpy's path_stem + ".npy" for fmt="npy"); returns the file name."""
    if fmt == "h5":
        from . import h5

        path = path_stem + ".h5"
        h5.write_dataset(path, "steps", np.asarray(steps, dtype=np.uint32))
    elif fmt == "npy":
        path = path_stem + ".npy"
        np.save(path, np.asarray(steps, dtype=np.uint32))
    else:
        raise ValueError("fmt must be 'h5' or 'npy'")
    return path


def run_scan(configs, num_points=100, num_iterations=int(1e8), out_dir=".", mode="exact", table_budget=64 << 30,
             save=True, fmt="h5"):
    """Tracks every configuration as in run_sim.py and returns {config: steps}.
    Configurations sharing (omega_x, omega_y, epsilon) share one modulation
    table; jobs are batched while their distinct tables fit in table_budget
    bytes (32 bytes per turn each)."""
    from . import henon_tracker, particles, track_batch

    if save and fmt == "h5":
        from . import h5

        h5.lib()  # no HDF5 library: fail before tracking
    configs = list(configs)
    per_table = 32 * int(num_iterations)
    if per_table > table_budget:
        raise ValueError("one modulation table exceeds the table budget")
    results = {}
    trackers = {}
    batch = []

    def flush():
        if not batch:
            return
        track_batch([(trackers[(c.omega_x, c.omega_y, c.epsilon)], p, num_iterations, c.mu, 10.0)
                     for c, p in batch], mode)
        for c, p in batch:
            steps = p.get_steps().astype(np.uint32)
            results[c] = steps
            if save:
                save_steps(os.path.join(out_dir, c.file_stem()), steps, fmt)
        batch.clear()
        trackers.clear()

    for c in configs:
        key = (c.omega_x, c.omega_y, c.epsilon)
        if key not in trackers and (len(trackers) + 1) * per_table > table_budget:
            flush()
        if key not in trackers:
            trackers[key] = henon_tracker(int(num_iterations), c.omega_x, c.omega_y, "sps", epsilon=c.epsilon)
        batch.append((c, particles(*initial_conditions(c.max_radius, num_points))))
    flush()
    return results


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    # htcondor_jobs/run_sim.py:15-24
    parser.add_argument("--omegax", type=float, default=0.31)
    parser.add_argument("--omegay", type=float, default=0.32)
    parser.add_argument("--epsilon", type=float, default=0.0)
    parser.add_argument("--mu", type=float, default=0.0)
    parser.add_argument("--max_radius", type=float, default=0.25)
    parser.add_argument("--num_points", type=int, default=100)
    parser.add_argument("--num_iterations", type=int, default=int(1e8))
    parser.add_argument("--jobs", default=None, help="job list (compose_config's format): run every line")
    parser.add_argument("--out_dir", default=".")
    parser.add_argument("--mode", default="exact", choices=("exact", "fast"))
    parser.add_argument("--format", default="h5", choices=("h5", "npy"))
    args = parser.parse_args(argv)
    if args.jobs:
        configs = read_config(args.jobs)
    else:
        configs = [ScanConfig(args.omegax, args.omegay, args.epsilon, args.mu, args.max_radius)]
    start = datetime.datetime.now()
    run_scan(configs, args.num_points, args.num_iterations, args.out_dir, args.mode, fmt=args.format)
    print("Elapsed time: {}".format(datetime.datetime.now() - start))


if __name__ == "__main__":
    main()
