"""Writes tests/golden/reference_vectors.npz: inputs and outputs of the reference's
own CPU code -- build_graph (graph.h:190-199), Graph::search (graph.h:33-115) and
BitHash (bitHash.h:17-56) -- compiled from its sources into oracle/_ref
(make -C oracle).  The fixture lets the build and the oracle be checked against the
reference where its sources are absent.  Run in the development container:
python tests/golden/make_ref_vectors.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref is not built (make -C oracle with the reference checkout present)")
    rng = np.random.default_rng(2024)
    # integer data (distance ties) and float data
    int_base = rng.integers(0, 64, size=(600, 12)).astype(np.float32)
    int_q = rng.integers(0, 64, size=(16, 12)).astype(np.float32)
    flt_base = rng.random((500, 10), dtype=np.float32)
    flt_q = rng.random((16, 10), dtype=np.float32)
    out = {"int_base": int_base, "int_queries": int_q, "float_base": flt_base, "float_queries": flt_q}
    for tag, base, q in (("int", int_base, int_q), ("float", flt_base, flt_q)):
        adj = oracle.ref_build_graph(base)
        out[f"{tag}_adj"] = adj
        for ef in (10, 50):
            ids, tr, _, _ = oracle.ref_graph_search(base, adj, q, 10, ef)
            out[f"{tag}_ids_ef{ef}"] = ids  # worst-first, as Graph::search returns them
            out[f"{tag}_trace_ef{ef}"] = tr
    vecs = rng.normal(size=(40, 24)).astype(np.float32)
    out["bithash_vectors"] = vecs
    out["bithash_matrix"] = oracle.ref_bithash_matrix(24, 42)
    out["bithash_codes"] = oracle.ref_bithash(vecs, 42)  # 8 x uint32 per vector
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)


if __name__ == "__main__":
    main()
