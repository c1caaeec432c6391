"""Writes tests/golden/ed_energies.json: exact ground energies used to pin the oracle.

Run from the repository root:  python tests/golden/make_golden.py
Energies come from sparse Lanczos (scipy ``eigsh``) on the fixed-Sz sector of the
Hamiltonian (oracle/ed.py); they are independent of every tensor-network routine.
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ed  # noqa: E402


def build_tree_edges(nbranch=3, nbranch_sites=3):
    """Vertices/edges of test/utilities/tree_graphs.jl:10-23 (centre (0,0), branches)."""
    verts = [(0, 0)] + [(b, s) for b in range(1, nbranch + 1) for s in range(1, nbranch_sites + 1)]
    edges = []
    for b in range(1, nbranch + 1):
        edges.append(((0, 0), (b, 1)))
        for s in range(2, nbranch_sites + 1):
            edges.append(((b, s - 1), (b, s)))
    return verts, edges


def main():
    out = {"heisenberg_s12_chain": {}, "xxz_s12_chain": {}, "heisenberg_s1_chain": {}, "tree": {}}
    for N in (4, 6, 8, 10, 12, 14, 16, 20):
        out["heisenberg_s12_chain"][str(N)] = ed.xxz_chain_ground_energy(N)
    for N, jz in ((10, 0.0), (10, 0.5), (10, 1.5), (12, 2.0), (16, 0.3)):
        out["xxz_s12_chain"][f"{N}_{jz}"] = ed.xxz_chain_ground_energy(N, Jz=jz)
    out["heisenberg_s1_chain"]["10"] = ed.xxz_chain_ground_energy(10, d=3)
    verts, edges = build_tree_edges()
    idx = {v: i for i, v in enumerate(verts)}
    H, _ = ed.xxz_hamiltonian(len(verts), [(idx[a], idx[b]) for a, b in edges], m2_total=0)
    out["tree"]["build_tree_3_3_heisenberg_sz0"] = ed.ground_energy(H)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ed_energies.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
