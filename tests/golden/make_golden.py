"""Writes the golden fixtures in this directory: small seeded graphs as .bin
edge lists plus the labels the oracle (oracle/liboracle.so) assigns at every
step, as .npz (no pickles).  Run from the repository root:

    python tests/golden/make_golden.py

The oracle is pinned by tests/test_oracle.py (closed forms, reductions,
known-answer vectors, the reference's own loader and tools); these files
freeze its output so both the oracle and the GPU path are checked against
data that does not change when code changes.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import glp_amd as G  # noqa: E402
import oracle_lib as O  # noqa: E402


def save_graph(name, g, weighted=False):
    path = os.path.join(HERE, name + ".bin")
    g.save_bin(path)
    # the fixture is what the reader makes of the file
    return O.load_bin(path, weighted=weighted)


def main():
    # planted partition, 4096 vertices, 8 communities (a scaled-down 65,536 config)
    g = G.Graph.planted(4096, 8, 12.0, 4.0, seed=1)
    off, nbr, _ = save_graph("planted_4096", g)
    np.savez_compressed(os.path.join(HERE, "planted_4096.npz"), classic=O.classic(off, nbr, 20))

    # R-MAT scale 11, degree-sorted: classic, layered (three gammas)
    g = G.Graph.rmat(11, 16, seed=7)
    off, nbr, _ = save_graph("rmat_s11", g)
    lay = {"layered_%d_%d" % pq: O.layered(off, nbr, 6, *pq)[0] for pq in [(0, 1), (1, 4), (1, 1)]}
    np.savez_compressed(os.path.join(HERE, "rmat_s11.npz"), classic=O.classic(off, nbr, 8), **lay)

    # weighted R-MAT scale 10 (u, v, w) triples
    g = G.Graph.rmat(10, 16, seed=8, wmax=8)
    off, nbr, w = save_graph("rmat_s10_w", g, weighted=True)
    np.savez_compressed(os.path.join(HERE, "rmat_s10_w.npz"), classic=O.classic(off, nbr, 6, w=w),
                        layered_1_4=O.layered(off, nbr, 5, 1, 4, w=w)[0])

    # LFR-style graph: SLPA memory rows and post-processing
    g = G.Graph.lfr(2048, seed=9)
    off, nbr, _ = save_graph("lfr_2048", g)
    mem = O.slpa(off, nbr, seed=7, T=10)
    dom, ns, sets = O.slpa_post(mem, 1, 10, 8)
    np.savez_compressed(os.path.join(HERE, "lfr_2048.npz"), slpa_mem=mem, dominant=dom, nsets=ns, sets=sets)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
