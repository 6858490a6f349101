"""Pins the oracle's generator (load_circuit, circuit.cc:1071-1288).

Independent evidence used here:
  * the POSIX drand48 definition, implemented with Python integers, against
    the glibc stream the oracle draws from;
  * a literal model of load_circuit over Python ``set`` objects (sorted-set
    walk for random_element, erase/insert on the privacy maps, find_location
    by membership) on tiny circuits;
  * draw-count and classification invariants at the 4-piece size (4 x 1024 nodes x 4096 wires);
  * the committed digest of that circuit (tests/golden/generator_4x1024x4096.json).
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

A, INC, M = 0x5DEECE66D, 0xB, 1 << 48


def posix_drand48(seed, n):
    x = ((seed & 0xFFFFFFFF) << 16) | 0x330E
    out = []
    for _ in range(n):
        x = (A * x + INC) % M
        out.append(x / float(M))
    return out


@pytest.mark.parametrize("seed", [12345, 0, 1, -1, 2**31 - 1, -123456])
def test_drand48_matches_posix_definition(O, seed):
    got = O.drand48_stream(seed, 2000)
    want = np.array(posix_drand48(seed, 2000))
    assert np.array_equal(got, want)


class LiteralGenerator:
    """load_circuit restated over Python sets, for tiny circuits."""

    def __init__(self, seed):
        self.u = iter(posix_drand48(seed, 100000))
        self.draws = 0

    def d(self):
        self.draws += 1
        return next(self.u)

    def random_element(self, s):
        index = int(self.d() * len(s))
        it = iter(sorted(s))
        v = next(it)
        while index > 0:
            v = next(it)
            index -= 1
        return v

    def run(self, P, npp, wpp, pct, stable=False):
        f32 = np.float32
        nodes, wires = [], []
        private_node_map = [set() for _ in range(P)]
        shared_node_map = [set() for _ in range(P)]
        ghost_node_map = [set() for _ in range(P)]
        privacy = [set(), set()]
        ptr = 0
        for n in range(P):
            for _ in range(npp):
                v = f32(2 * self.d() - 1)
                c = f32(self.d() + 1)
                leak = f32(float(f32(0.1)) * self.d())
                nodes.append((0.0, v, c, leak))
                private_node_map[n].add(ptr)
                privacy[0].add(ptr)
                ptr += 1
        for n in range(P):
            for _ in range(wpp):
                if stable:
                    R = f32(self.d() * 10 + 1)
                    L = f32((self.d() * 0.01 + 0.1) * 1e-6)
                    C = f32(self.d() * 0.01 + 0.001)
                else:
                    R = f32(self.d() * 10 + 1)
                    L = f32(self.d() * 0.01 + 0.1)
                    C = f32(self.d() * 0.1)
                inp = self.random_element(private_node_map[n])
                if 100 * self.d() < pct:
                    out = self.random_element(private_node_map[n])
                else:
                    nn = int(self.d() * (P - 1))
                    if nn >= n:
                        nn += 1
                    out = self.random_element(private_node_map[nn])
                    privacy[0].discard(out)
                    privacy[1].add(out)
                    ghost_node_map[n].add(out)
                wires.append([n, inp, out, L, R, C])
        locator = []
        for n in range(P):
            for i in range(n * npp, (n + 1) * npp):
                if i not in privacy[0]:
                    private_node_map[n].discard(i)
                    shared_node_map[n].add(i)
                    locator.append(0)
                else:
                    locator.append(1)

        def find_location(p, n):
            if p in private_node_map[n]:
                return 0
            if p in shared_node_map[n]:
                return 1
            if p in ghost_node_map[n]:
                return 2
            raise AssertionError("unreachable")

        out_w = []
        for n, inp, out, L, R, C in wires:
            out_w.append((inp, out, find_location(inp, n), find_location(out, n), L, R, C))
        return nodes, out_w, locator


@pytest.mark.parametrize("P,npp,wpp,pct,seed,stable", [
    (2, 4, 8, 50, 3, False), (3, 5, 7, 60, 12345, False), (4, 3, 6, 0, 99, True),
    (2, 1, 5, 50, 5, False), (5, 6, 4, 95, 2024, True)])
def test_generator_matches_literal_set_model(O, P, npp, wpp, pct, seed, stable):
    prof = O.PROFILE_STABLE if stable else O.PROFILE_REFERENCE
    c = O.generate(P, npp, wpp, pct, seed, prof)
    lit = LiteralGenerator(seed)
    nodes, wires, loc = lit.run(P, npp, wpp, pct, stable)
    assert c.draws == lit.draws
    for i, (q, v, cap, leak) in enumerate(nodes):
        got = c.nodes[i]
        assert (got["charge"], got["voltage"], got["capacitance"], got["leakage"]) == (q, v, cap, leak)
    for k, (inp, out, il, ol, L, R, C) in enumerate(wires):
        w = c.wires[k]
        assert (w["in_ptr"], w["out_ptr"], w["in_loc"], w["out_loc"]) == (inp, out, il, ol)
        assert (w["inductance"], w["resistance"], w["capacitance"]) == (L, R, C)
        assert not w["current"].any() and not w["voltage"].any()
    assert list(c.locator) == loc


def _classification_invariants(c, P, npp, wpp):
    w = c.wires
    piece_of_wire = np.arange(len(w)) // wpp
    in_piece = w["in_ptr"] // npp
    out_piece = w["out_ptr"] // npp
    cross = out_piece != piece_of_wire
    # in-node always in the wire's own piece (circuit.cc:1166)
    assert np.all(in_piece == piece_of_wire)
    shared = c.locator == 0
    # every shared node is the out-end of at least one cross wire, and the
    # out-end of every cross wire is shared (circuit.cc:1183-1184,1210-1216)
    targets = np.zeros(len(c.nodes), bool)
    targets[w["out_ptr"][cross]] = True
    assert np.array_equal(shared, targets)
    # locations: ghost iff the end is in another piece; else shared iff the
    # node is shared (find_location order, circuit.cc:1054-1065)
    for end, piece in (("in", in_piece), ("out", out_piece)):
        loc = w[f"{end}_loc"]
        ptr = w[f"{end}_ptr"]
        other = piece != piece_of_wire
        assert np.all(loc[other] == 2)
        assert np.all(loc[~other] == np.where(shared[ptr[~other]], 1, 0))
    return cross


def test_small_draw_count_and_classification(O):
    P, npp, wpp = 4, 1024, 4096
    c = O.generate(P, npp, wpp, 95, 12345)
    cross = _classification_invariants(c, P, npp, wpp)
    # 3 draws per node; 6 per in-piece wire, 7 per cross-piece wire
    assert c.draws == 3 * P * npp + 6 * P * wpp + int(cross.sum())
    # shared fraction ~ 1 - exp(-cross ends / nodes) for uniform targets
    expect = 1 - math.exp(-cross.sum() / (P * npp))
    assert abs((c.locator == 0).mean() - expect) < 0.02
    # value ranges of circuit.cc:1123-1125,1162-1164
    n = c.nodes
    assert np.all((n["voltage"] >= -1) & (n["voltage"] <= 1))
    assert np.all((n["capacitance"] >= 1) & (n["capacitance"] <= 2))
    assert np.all((n["leakage"] >= 0) & (n["leakage"] <= np.float32(0.1)))
    w = c.wires
    assert np.all((w["resistance"] >= 1) & (w["resistance"] <= 11))
    assert np.all((w["inductance"] >= np.float32(0.1)) & (w["inductance"] <= np.float32(0.11)))
    assert np.all((w["capacitance"] >= 0) & (w["capacitance"] <= np.float32(0.1)))


def test_pct50_classification(O):
    c = O.generate(4, 1024, 4096, 50, 12345)
    cross = _classification_invariants(c, 4, 1024, 4096)
    expect = 1 - math.exp(-cross.sum() / 4096)
    assert abs((c.locator == 0).mean() - expect) < 0.03


def test_single_piece_cross_branch_wraps_to_own_piece(O):
    # With one piece the reference would index piece 1; the pick wraps to
    # piece 0 and the target node becomes shared.
    c = O.generate(1, 64, 256, 50, 12345)
    w = c.wires
    assert np.all(w["out_ptr"] < 64)
    assert (c.locator == 0).any()
    assert np.all(w["in_loc"] != 2) and np.all(w["out_loc"] != 2)


def _digest(c):
    h = {}
    for name in ("nodes", "wires", "locator"):
        h[name] = hashlib.sha256(getattr(c, name).tobytes()).hexdigest()
    return h


def test_small_golden_digest(O):
    with open(os.path.join(GOLDEN, "generator_4x1024x4096.json")) as f:
        g = json.load(f)
    c = O.generate(4, 1024, 4096, 95, 12345)
    assert _digest(c) == g["sha256"]
    assert c.draws == g["draws"]
    for k, v in g["first_nodes"].items():
        assert np.array_equal(c.nodes[int(k)].tolist(), v)
