"""Writes the committed golden fixtures under tests/golden/.

  graphs/<name>.npz        graph files the reference ships in test/test_graphs/
                           (data, converted to a compact form: a packed upper
                           triangle for dense graphs, an edge list otherwise)
  known_answers.json       their exact MVC from the reference's own table
                           (test/runGraphs.py:8-30), with edge checksums
  oracle_vectors.json      seeded generated graphs with U (RemoveMax, seeded
                           RemoveEdge), MVC, node and rule-call counts and PVC
                           answers, computed by the oracle restatement AND, when
                           the reference sources are present, by the reference's
                           own CPU code (oracle/_ref/ref_driver); the script
                           stops if the two disagree.

Run from the repository root:  python tests/golden/make_golden.py [--with-slow]
(--with-slow adds the 200-vertex Erdos-Renyi workload graph, ~5 min of CPU;
--only-known rewrites graphs/ and known_answers.json only).
"""
import argparse
import io
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "parallel-vertex-cover-algorithms-on-gpus_amd"))
import graphgen  # noqa: E402
import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
REF_GRAPHS = "/root/reference/test/test_graphs"

# Known answers of test/runGraphs.py:8-30 for graphs shipped in the reference
# and small enough to keep as fixtures.
KNOWN = {
    "cop_hat300-1.txt": 292,
    "cop_hat300-2.txt": 275,
    "cop_hat500-1.txt": 491,
    "cop_hat700-1.txt": 689,
    "cop_hat1000-1.txt": 990,
    "Sister-Cities.txt": 5527,
    "formated_maayan-figeys.txt": 316,
    "formated_maayan-vidal.txt": 1077,
    "formated_movielens-100k_rating.txt": 849,
    "formated_wikipedia_link_csb.txt": 1998,
    "lastfm-asia.txt": 3447,
}
# Shipped with an answer in the same table but too hard for the CPU oracle to
# confirm in minutes; kept as data for GPU runs and marked "cpu_confirmed": false.
KNOWN_HARD = {
    "cop_hat300-3.txt": 264,
    "cop_hat500-2.txt": 464,
    "cop_hat500-3.txt": 450,
    "cop_hat700-2.txt": 656,
    "cop_hat1000-2.txt": 954,
    "Power-Grid.txt": 2203,
    "vc-exact_009.txt": 21348,
    "vc-exact_023.txt": 16013,
    "formated_wikipedia_link_lo.txt": 1536,
}
# Files whose vertices are numbered from 0 (the reference's loader,
# src/auxFunctions.cpp:151-153, expects 1 and cannot read them).
ZERO_INDEXED = {
    "formated_maayan-figeys.txt", "formated_maayan-vidal.txt", "formated_movielens-100k_rating.txt",
    "formated_wikipedia_link_csb.txt", "formated_wikipedia_link_lo.txt", "lastfm-asia.txt",
}


def load_reference_graph(path, index_base):
    """A reference graph file as an oracle Graph: the oracle's loader for the
    reference's 1-indexed format, a plain token parse for 0-indexed files."""
    if index_base == 1:
        return O.load(path)
    tok = np.array(open(path).read().split(), dtype=np.int64)
    n, m = int(tok[0]), int(tok[1])
    edges = tok[2:2 + 2 * m].reshape(m, 2) - index_base
    return O.from_edges(n, edges)


def pack_graph(n, edges):
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    lo = np.minimum(edges[:, 0], edges[:, 1])
    hi = np.maximum(edges[:, 0], edges[:, 1])
    dense_bits = n * (n - 1) // 2
    buf_e = io.BytesIO()
    dt = np.uint16 if n < 65536 else np.uint32
    np.savez_compressed(buf_e, n=np.array([n]), edges=np.stack([lo, hi], 1).astype(dt))
    if dense_bits <= 8_000_000:
        idx = lo * n - lo * (lo + 1) // 2 + (hi - lo - 1)
        bits = np.zeros(dense_bits, dtype=np.uint8)
        bits[idx] = 1
        buf_t = io.BytesIO()
        np.savez_compressed(buf_t, n=np.array([n]), triu=np.packbits(bits))
        if len(buf_t.getvalue()) < len(buf_e.getvalue()):
            return buf_t.getvalue()
    return buf_e.getvalue()


def unpack_graph(path):
    z = np.load(path, allow_pickle=False)
    n = int(z["n"][0])
    if "edges" in z:
        return n, z["edges"].astype(np.uint32)
    bits = np.unpackbits(z["triu"])[: n * (n - 1) // 2]
    iu, ju = np.triu_indices(n, 1)
    keep = bits.astype(bool)
    return n, np.stack([iu[keep], ju[keep]], 1).astype(np.uint32)


def reference_fixtures():
    out = {}
    os.makedirs(os.path.join(HERE, "graphs"), exist_ok=True)
    for fname, mvc in list(KNOWN.items()) + list(KNOWN_HARD.items()):
        path = os.path.join(REF_GRAPHS, fname)
        base = 0 if fname in ZERO_INDEXED else 1
        g = load_reference_graph(path, base)
        assert O.check_graph(g), fname
        src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.srcPtr.astype(np.int64)))
        keep = src < g.dst
        edges = np.stack([src[keep], g.dst[keep]], 1)
        name = os.path.splitext(fname)[0]
        blob = pack_graph(g.n, edges)
        with open(os.path.join(HERE, "graphs", name + ".npz"), "wb") as f:
            f.write(blob)
        n2, e2 = unpack_graph(os.path.join(HERE, "graphs", name + ".npz"))
        assert n2 == g.n and graphgen.edge_checksum(e2) == graphgen.edge_checksum(edges)
        entry = {"file": f"test/test_graphs/{fname}", "index_base": base, "n": int(g.n), "m": int(g.m), "mvc": mvc,
                 "source": "test/runGraphs.py:8-30", "edge_sha256": graphgen.edge_checksum(edges),
                 "cpu_confirmed": fname in KNOWN}
        if fname in KNOWN:
            U = O.remove_max(g)
            r, st, _ = O.sequential(g, U)
            assert r == mvc, (fname, r, mvc)
            entry["U_max"] = U
            entry["oracle_nodes"] = st["nodes"]
        out[name] = entry
        print(name, entry, flush=True)
    return out


def generated_cases(with_slow):
    cases = []
    for s in range(12):
        cases.append(("er", {"n": 60 + 5 * s, "p": 0.06, "seed": 10 + s}))
    for s in range(6):
        cases.append(("er", {"n": 40, "p": 0.2, "seed": 50 + s}))
    for s in range(6):
        cases.append(("ba", {"n": 300 + 50 * s, "m": 2 + s % 3, "seed": 70 + s}))
    for s in range(8):
        cases.append(("reg3", {"n": 120, "seed": s}))
    cases.append(("single", {"name": "single_ba"}))
    cases.append(("single", {"name": "single_reg3"}))
    if with_slow:
        cases.append(("single", {"name": "single_er"}))
    return cases


def make_graph(kind, spec):
    if kind == "er":
        return spec["n"], graphgen.erdos_renyi(spec["n"], spec["p"], spec["seed"])
    if kind == "ba":
        return spec["n"], graphgen.barabasi_albert(spec["n"], spec["m"], spec["seed"])
    if kind == "reg3":
        return spec["n"], graphgen.regular3_rewired(spec["n"], spec["seed"])
    return graphgen.single(spec["name"])


EDGE_SEED = 12345


def oracle_vectors(with_slow, old):
    out = []
    use_ref = O.ref_available()
    for kind, spec in generated_cases(with_slow):
        n, e = make_graph(kind, spec)
        key = json.dumps([kind, spec], sort_keys=True)
        cs = graphgen.edge_checksum(e)
        prev = old.get(key)
        if prev is not None and prev["edge_sha256"] == cs:
            out.append(prev)
            continue
        g = O.from_edges(n, e)
        U = O.remove_max(g)
        Ue = O.remove_edge(g, EDGE_SEED)
        mvc, st, cover = O.sequential(g, U)
        rec = {"key": key, "kind": kind, "spec": spec, "n": n, "m": int(len(e)), "edge_sha256": cs, "U_max": U,
               "U_edge_seed": EDGE_SEED, "U_edge": Ue, "mvc": mvc,
               "oracle": {k: st[k] for k in ("nodes", "leaf_calls", "tri_calls", "high_calls", "square_calls",
                                             "delete_calls")},
               "pvc": {}}
        for k in (mvc - 1, mvc, mvc + 1):
            if k < 0:
                continue
            kf, pst, _ = O.sequential_parameterized(g, U, k)
            rec["pvc"][str(k)] = {"kFound": kf, "nodes": pst["nodes"]}
        if use_ref:
            with tempfile.TemporaryDirectory() as td:
                path = os.path.join(td, "g.txt")
                graphgen.write_graph_file(path, n, e)
                ref = O.ref_run(path, seed=EDGE_SEED, timeout=3600)
                for a, b in (("U_max", U), ("U_edge", Ue), ("result", mvc), ("nodes", st["nodes"]),
                             ("leaf_calls", st["leaf_calls"]), ("tri_calls", st["tri_calls"]),
                             ("high_calls", st["high_calls"]), ("square_calls", st["square_calls"]),
                             ("delete_calls", st["delete_calls"])):
                    assert ref[a] == b, (key, a, ref[a], b)
                for k, v in rec["pvc"].items():
                    refp = O.ref_run(path, pvc_k=int(k), timeout=3600)
                    assert refp["kFound"] == v["kFound"] and refp["nodes"] == v["nodes"], (key, k, refp, v)
            rec["ref_checked"] = True
        else:
            rec["ref_checked"] = False
        print(kind, spec, "U", U, "mvc", mvc, "nodes", st["nodes"], "ref", rec["ref_checked"], flush=True)
        out.append(rec)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--with-slow", action="store_true")
    ap.add_argument("--skip-known", action="store_true")
    ap.add_argument("--only-known", action="store_true")
    args = ap.parse_args()
    if not args.skip_known:
        if not os.path.isdir(REF_GRAPHS):
            sys.exit("reference graphs not present; use --skip-known")
        known = reference_fixtures()
        with open(os.path.join(HERE, "known_answers.json"), "w") as f:
            json.dump(known, f, indent=1, sort_keys=True)
    if args.only_known:
        return
    vec_path = os.path.join(HERE, "oracle_vectors.json")
    old = {}
    if os.path.exists(vec_path):
        for r in json.load(open(vec_path)):
            old[r["key"]] = r
    vecs = oracle_vectors(args.with_slow, old)
    if not args.with_slow:
        vecs += [r for k, r in old.items() if r["kind"] == "single" and r["spec"].get("name") == "single_er"]
    with open(vec_path, "w") as f:
        json.dump(vecs, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
