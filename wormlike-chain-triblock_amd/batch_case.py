"""Heterogeneous SCFT instances of the batching workload.

48 independent linear ABC wormlike triblocks from seed 3: grid edge uniform over
{24, 32, 48, 64} (cubic grid), Lmax over {4, 6} (every block), f_A, f_B ~ U[0.2, 0.45],
f_C = 1 - f_A - f_B, chiN_AB, chiN_BC, chiN_AC ~ U[12, 30], lambda ~ U[0.25, 4] (every
block), box edge ~ U[3, 6] (cubic box); initial w = 0.1 N(0, 1) per block.  Ns is not part
of the draw; it is fixed at 100 contour steps, split NINT(f Ns) (initialize_mod.f90:142-144).
Draw order per instance: edge, Lmax, f_A, f_B, chi_AB, chi_BC, chi_AC, lambda, box, then
the fields ([3][n][n][n], x fastest), all from one numpy PCG64 stream.
"""
import numpy as np

EDGES = (24, 32, 48, 64)
LMAX = (4, 6)
NS = 100


def block_lengths(Ns, fracs):
    out = [int(np.floor(f * Ns + 0.5)) for f in fracs[:-1]]
    return tuple(out + [Ns - sum(out)])


def draw(n=48, seed=3, Ns=NS):
    """List of instance dicts: shape, box, Nblk, L, lam, chi (3x3 symmetric), f, w."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        e = int(EDGES[rng.integers(len(EDGES))])
        L = int(LMAX[rng.integers(len(LMAX))])
        fA = float(rng.uniform(0.2, 0.45))
        fB = float(rng.uniform(0.2, 0.45))
        fC = 1.0 - fA - fB
        cab, cbc, cac = (float(v) for v in rng.uniform(12.0, 30.0, 3))
        lam = float(rng.uniform(0.25, 4.0))
        box = float(rng.uniform(3.0, 6.0))
        w = 0.1 * rng.standard_normal((3, e, e, e))
        chi = np.array([[0.0, cab, cac], [cab, 0.0, cbc], [cac, cbc, 0.0]])
        out.append(dict(shape=(e, e, e), box=(box,) * 3, Nblk=block_lengths(Ns, (fA, fB, fC)),
                        L=(L,) * 3, lam=(lam,) * 3, chi=chi, f=(fA, fB, fC), w=w))
    return out
