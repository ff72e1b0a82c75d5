"""Seeded synthetic workloads of the batched-apply and serving measurements.

Both the benchmark and the GPU tests draw their inputs here, so that a test
exercises exactly the workload a benchmark figure is quoted on.
"""
import numpy as np


def batch_workload(seed=4, count=64):
    """Uneven batch of Helmholtz applies: hex/tet alternating (50/50), P ~
    U{3..9} (nummodes 4..10), Q = P+2 in direction 0 (tets: Q-1 Gauss-Radau
    points in directions 1 and 2), element counts log-uniform in [1e3, 1e5],
    affine elements of size ~0.1 with a per-element scale U[0.5, 2], lambda
    ~ U[0.1, 10].  Returns [(shape "Hex"|"Tet", nummodes, numpoints0, E, jac,
    df, lambda)] with per-element jac[e] and df[e*9 + j*3 + k] = dxi_k/dx_j."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        shape = "Hex" if i % 2 == 0 else "Tet"
        nm = int(rng.integers(4, 11))
        E = int(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        s = rng.uniform(0.5, 2.0, E)[:, None, None]
        A = np.eye(3)[None] + 0.25 * rng.uniform(-1, 1, (E, 3, 3))
        M = s * A * 0.05  # d x / d xi of a hex; a tet's map takes M / 2
        jac = np.abs(np.linalg.det(M)) / (1.0 if shape == "Hex" else 8.0)
        Minv = np.linalg.inv(M if shape == "Hex" else M / 2)
        df = np.ascontiguousarray(np.transpose(Minv, (0, 2, 1))).reshape(-1)
        out.append((shape, nm, nm + 1, E, jac, df, float(rng.uniform(0.1, 10))))
    return out


def serve_requests(count, seed=5):
    """Helmholtz solve requests: hex box with E log-uniform in [2^10, 2^17]
    elements (n = round(E^(1/3)) per side), P ~ U{3..8}, Q = P+2, lambda ~
    U[0.1, 10], forcing frequency k ~ U{1, 2, 3} (manufactured sine
    solution), tolerance 1e-10."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        E = float(np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 17))))
        P = int(rng.integers(3, 9))
        out.append(dict(n=max(1, int(round(E ** (1 / 3)))), nummodes=P + 1, numpoints=P + 2,
                        lam=float(rng.uniform(0.1, 10.0)), k=int(rng.integers(1, 4)), tol=1e-10, maxit=5000))
    return out


def poisson_arrivals(count, rate, seed=55):
    """Arrival times (seconds, first at 0) of a Poisson stream of `rate`
    requests per second."""
    arr = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, count))
    return arr - arr[0]
