"""Seeded initial conditions and the reference's restart-file format.

The reference starts from rest with T = 0 (src/module_solve_thermal.f90:81)
and takes other initial states only through its restart files
(mpi_Post_FileIn_Continue_Post_Reassembly_IO, src/module_post.f90:1264-1357):
cont_time.bin holds (time, dt) as two FP64, cont_{U,V,W,P,THETA}.bin hold
the n1m*n2m*n3m interior as raw little-endian FP64 in Fortran order.
"""
import os

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_uniform(seed, n):
    """n doubles in [0,1): splitmix64 stream, top 53 bits of each output."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cell_centres(x):
    """y_c(j) = (x(j) + x(j+1))/2 for j = 1..nm from face coordinates x(0:nm+1)."""
    return 0.5 * (x[1:-1] + x[2:])


def conduction_noise_T(cfg, x2, amp, seed):
    """Interior T (n3m, n2m, n1m) C-order = Fortran (n1m,n2m,n3m):
    Tbt + (Tup - Tbt) * y_c / L2 + amp * (2u - 1), u ~ splitmix64(seed)."""
    yc = cell_centres(np.asarray(x2))
    base = cfg.Tbt + (cfg.Tup - cfg.Tbt) * yc / cfg.L2
    u = splitmix64_uniform(seed, cfg.n1m * cfg.n2m * cfg.n3m).reshape(cfg.n3m, cfg.n2m, cfg.n1m)
    return base[None, :, None] + amp * (2.0 * u - 1.0)


def embed_interior(interior):
    """Interior (n3m,n2m,n1m) -> full (n3m+2,n2m+2,n1m+2) with zero ghosts."""
    n3, n2, n1 = interior.shape
    full = np.zeros((n3 + 2, n2 + 2, n1 + 2))
    full[1:-1, 1:-1, 1:-1] = interior
    return full


def write_restart(directory, time, dt, fields):
    """Write cont_time.bin and cont_{U,V,W,P,THETA}.bin (interiors, Fortran order)."""
    os.makedirs(directory, exist_ok=True)
    np.array([time, dt], dtype="<f8").tofile(os.path.join(directory, "cont_time.bin"))
    for name in ("U", "V", "W", "P", "THETA"):
        np.ascontiguousarray(fields[name], dtype="<f8").tofile(os.path.join(directory, f"cont_{name}.bin"))


def read_restart(directory, n1m, n2m, n3m):
    time, dt = np.fromfile(os.path.join(directory, "cont_time.bin"), dtype="<f8")[:2]
    out = {}
    for name in ("U", "V", "W", "P", "THETA"):
        a = np.fromfile(os.path.join(directory, f"cont_{name}.bin"), dtype="<f8")
        out[name] = a.reshape(n3m, n2m, n1m)
    return float(time), float(dt), out
