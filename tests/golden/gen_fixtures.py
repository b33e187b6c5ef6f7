"""Writes the golden fixtures (tests/golden/*.npz) from the independent numpy
restatement np_restatement.py.  Run: python tests/golden/gen_fixtures.py

Each fixture holds the inputs (grid, box, blocks, Lmax, lambda, w) and outputs
(q at s = Ns and a middle s, q* at s = 0, phi, Q, totden) of one tiny chain, plus
quadrature-built R matrices for Lmax = 8.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import np_restatement as R  # noqa: E402

CASES = {
    # AB diblock, even block lengths, different Lmax per block
    "ab_even": dict(shape=(4, 6, 8), box=(3.0, 2.5, 4.0), Nblk=(4, 6), L=(2, 3), lam=(1.0, 0.5),
                    seed=11, amp=1.5),
    # ABC triblock, odd block lengths, radix-3 edges, shrinking then growing Lmax
    "abc_odd_radix3": dict(shape=(6, 4, 6), box=(2.0, 3.0, 2.5), Nblk=(3, 5, 4), L=(1, 3, 2),
                           lam=(2.0, 0.3, 1.0), seed=12, amp=2.0),
    # one block, cubic grid
    "homopolymer": dict(shape=(8, 8, 8), box=(4.0, 4.0, 4.0), Nblk=(8,), L=(4,), lam=(0.7,),
                        seed=13, amp=1.0),
}


def make(name, c):
    rng = np.random.default_rng(c["seed"])
    w = c["amp"] * rng.standard_normal((len(c["Nblk"]),) + c["shape"])
    ch = R.Chain(c["shape"], c["box"], c["Nblk"], c["L"], c["lam"])
    q = ch.forward(w)
    qs = ch.backward(w)
    phi, Q, td = ch.density(q, qs)
    smid = ch.Ns // 2
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"), shape=np.array(c["shape"]), box=np.array(c["box"]),
        Nblk=np.array(c["Nblk"]), L=np.array(c["L"]), lam=np.array(c["lam"]), w=w,
        q_end=q[ch.Ns], q_mid=q[smid], s_mid=smid, qs_0=qs[0], phi=phi, Q=Q, totden=td)


def main():
    for name, c in CASES.items():
        make(name, c)
        print("wrote", name)
    np.savez_compressed(os.path.join(HERE, "rmatrices_L8.npz"), R=R.rmatrices_quadrature(8))
    print("wrote rmatrices_L8")


if __name__ == "__main__":
    main()
