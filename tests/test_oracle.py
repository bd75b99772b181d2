"""Pins the CPU oracle (oracle/si_oracle.cpp) against known answers, library
routines, closed forms and brute force -- not against a retyping of itself.

Reference behaviour restated by the oracle: src/pbwt/pbwt_index.cpp:20-247,
src/api/imputer.cpp:596-616,913-942, src/kernels/transition.cu:83-121,
src/kernels/emission.cu:35-54, src/kernels/forward_backward.cu:41-317,
src/kernels/logsumexp.cuh:67-70, src/kernels/sampling.cu:52-182.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, read_vcf_gt


def fixture_panel():
    pos, rows = read_vcf_gt(os.path.join(GOLDEN, "test_reference.vcf"))
    M = len(rows)
    ref = np.array([[a for g in r for a in g] for r in rows], np.uint8)  # haplotype 2s+p
    return np.array(pos, np.uint64), ref.reshape(M, -1)


# ----------------------------------------------------------------------- PBWT
def test_pbwt_fixture_known_answer():
    # Hand-derived from test_reference.vcf (H = 8): rs1 alleles by haplotype are
    # 0 0 0 1 1 0 1 1, so zeros (0,1,2,5) precede ones (3,4,6,7); rs2 then splits
    # that order into (0,1,2,3 | 5,4,6,7).
    _, ref = fixture_panel()
    prefix, div = oracle.pbwt_build(ref)
    assert ref.shape == (10, 8)
    assert prefix[0].tolist() == [0, 1, 2, 5, 3, 4, 6, 7]
    assert prefix[1].tolist() == [0, 1, 2, 3, 5, 4, 6, 7]
    for m in range(10):
        assert (div[m] == m).all()


def brute_prefix(ref):
    """Order after marker m = haplotypes sorted by the reversed prefix
    (x_m, x_{m-1}, ..., x_0) with nonzero alleles as 1, ties by index."""
    x = (ref != 0).astype(np.int64)
    M, H = x.shape
    out = np.empty((M, H), np.uint32)
    for m in range(M):
        keys = [np.arange(H)] + [x[r] for r in range(m + 1)]  # last key is primary
        out[m] = np.lexsort(keys)
    return out


@pytest.mark.parametrize("seed,M,H", [(1, 12, 7), (2, 25, 64), (3, 9, 65), (4, 30, 129), (5, 1, 5)])
def test_pbwt_brute_force(seed, M, H):
    rng = np.random.default_rng(seed)
    ref = (rng.random((M, H)) < 0.4).astype(np.uint8)
    ref[rng.random((M, H)) < 0.05] = 255  # missing alleles sort with the ones
    prefix, div = oracle.pbwt_build(ref)
    assert np.array_equal(prefix, brute_prefix(ref))
    assert all((div[m] == m).all() for m in range(M))
    for m in range(M):
        assert sorted(prefix[m].tolist()) == list(range(H))


@pytest.mark.parametrize("windows", [[1, 1, 28], [5, 7, 18], [30], [3, 0, 27]])
def test_pbwt_stream_equals_whole_build(windows):
    # the streamed sweep (window by window, order carried) gives the rows of
    # the whole-panel build, with global divergence indices
    rng = np.random.default_rng(sum(windows) + len(windows))
    ref = (rng.random((30, 77)) < 0.45).astype(np.uint8)
    ref[rng.random((30, 77)) < 0.03] = 255
    p, d = oracle.pbwt_build(ref)
    st = oracle.PbwtStream(77)
    m = 0
    for w in windows:
        pw, dw = st.next(ref[m:m + w])
        assert np.array_equal(pw, p[m:m + w]) and np.array_equal(dw, d[m:m + w])
        m += w
    assert st.m == 30 and np.array_equal(st.order, p[-1])


# ------------------------------------------------------------------ selection
def test_select_perm_known_answers():
    # Prefixes of the permutation libstdc++'s std::sort applies to H equal keys.
    assert oracle.select_perm(4432, 8).tolist() == [2952, 2960, 2959, 2958, 2957, 2956, 2955, 2954]
    assert oracle.select_perm(20000, 16).tolist() == [
        13330, 13339, 13338, 13337, 13336, 13335, 13334, 13333, 13332, 13331, 13340, 13329,
        13328, 13327, 13326, 13325]
    assert oracle.select_perm(8000, 8).tolist() == [5328, 5343, 5342, 5341, 5340, 5339, 5338, 5337]
    assert oracle.select_perm(40000, 8).tolist()[:2] == [26670, 26661]
    for H in range(1, 17):
        assert oracle.select_perm(H, 16).tolist() == list(range(H))


def test_select_states_fixture_marker0():
    # test_vcf_io.cpp:242-248 prints select_states(0, ..., 4) on the fixture:
    # H = 8 <= 16 keeps the prefix order, so the first four of prefix[0].
    _, ref = fixture_panel()
    prefix, div = oracle.pbwt_build(ref)
    assert oracle.select_states_sort(prefix[0], div[0], 4).tolist() == [0, 1, 2, 5]


@pytest.mark.parametrize("H,L", [(37, 8), (4432, 8), (300, 16), (5, 8), (3, 3)])
def test_state_table_equals_literal_sort(H, L):
    rng = np.random.default_rng(H)
    M = 6
    ref = (rng.random((M, H)) < 0.5).astype(np.uint8)
    prefix, div = oracle.pbwt_build(ref)
    h0, h1, a0, a1 = oracle.state_table(ref, prefix, L)
    for m in range(M):
        lit = oracle.select_states_sort(prefix[m], div[m], L)
        assert h0[m].tolist() == lit.tolist()
        assert h1[m].tolist() == [(h + 1) % H for h in lit]
        assert a0[m].tolist() == [ref[m, h] for h in lit]
        assert a1[m].tolist() == [ref[m, (h + 1) % H] for h in lit]


def test_state_table_short_panel_fill():
    # H < L: the first H come from the sort, the rest are i % H (pbwt_index.cpp:41-43)
    ref = np.array([[0, 1, 1]], np.uint8)
    prefix, div = oracle.pbwt_build(ref)
    h0, _, _, _ = oracle.state_table(ref, prefix, 8)
    assert h0[0].tolist() == [0, 1, 2, 0, 1, 2, 0, 1]
    assert oracle.select_states_sort(prefix[0], div[0], 8).tolist() == h0[0].tolist()


def test_state_table_simple_mode():
    ref = np.zeros((2, 8), np.uint8)
    h0, h1, _, _ = oracle.state_table(ref, np.zeros((2, 8), np.uint32), 3, simple_mode=True)
    assert h0[0].tolist() == [0, 2, 4] and h1[0].tolist() == [1, 3, 5]


# ------------------------------------------- the reference's GPU selection
def _pbwt_rows(H, M=5, seed=0):
    rng = np.random.default_rng(seed + H)
    ref = (rng.random((M, H)) < 0.4).astype(np.uint8)
    return ref, *oracle.pbwt_build(ref)


@pytest.mark.parametrize("H,L", [(128, 1), (128, 3), (200, 8), (4432, 8), (10000, 32), (640, 4)])
def test_gpu_warp_selection_closed_form(H, L):
    # every divergence of row m is m, so [p0, p32, p64, p96, 4, 5, ..., L-1]
    _, prefix, div = _pbwt_rows(H)
    for m in range(prefix.shape[0]):
        p = prefix[m]
        want = ([int(p[0]), int(p[32]), int(p[64]), int(p[96])] + list(range(4, L)))[:L]
        assert oracle.select_states_gpu(p, div[m], L, "warp").tolist() == want
        assert oracle.select_states_gpu(p, div[m], L).tolist() == want


@pytest.mark.parametrize("H,L", [(10001, 8), (20000, 16), (300, 33), (300, 64), (20, 40)])
def test_gpu_insert_selection_closed_form(H, L):
    # ties never displace: the first min(L, H) prefix positions, then i % H
    _, prefix, div = _pbwt_rows(H)
    for m in range(prefix.shape[0]):
        p = prefix[m]
        want = [int(p[i]) if i < H else i % H for i in range(L)]
        assert oracle.select_states_gpu(p, div[m], L, "insert").tolist() == want
        assert oracle.select_states_gpu(p, div[m], L).tolist() == want


@pytest.mark.parametrize("H,L", [(300, 65), (20000, 100), (40, 128), (33000, 128)])
def test_gpu_block_selection_closed_form(H, L):
    # L > 64 (the block kernel): ties never displace, so again the first
    # min(L, H) prefix positions, then i % H
    _, prefix, div = _pbwt_rows(H)
    for m in range(prefix.shape[0]):
        p = prefix[m]
        want = [int(p[i]) if i < H else i % H for i in range(L)]
        assert oracle.select_states_gpu(p, div[m], L, "block").tolist() == want
        assert oracle.select_states_gpu(p, div[m], L).tolist() == want


@pytest.mark.parametrize("H,L", [(256 * 80, 80), (256 * 100 + 17, 100)])
def test_gpu_block_selection_distinct_divergences_is_top_l(H, L):
    # with every thread's range holding at least L entries and distinct
    # divergences, the per-thread top-L lists and their merge give the global
    # top L by divergence, largest first (brute force: a full sort)
    rng = np.random.default_rng(H)
    div = rng.permutation(H * 3)[:H].astype(np.uint32)
    prefix = rng.permutation(H).astype(np.uint32)
    order = np.argsort(-div.astype(np.int64), kind="stable")[:L]
    assert oracle.select_states_gpu(prefix, div, L, "block").tolist() == prefix[order].tolist()


def test_gpu_warp_selection_hand_traced_rounds():
    p = np.arange(1000, 1160, dtype=np.uint32)  # prefix row: position i holds haplotype 1000+i
    # a larger divergence at lane 5's first entry wins only the first selection pass
    d = np.ones(128, np.uint32)
    d[5] = 100
    assert oracle.select_states_gpu(p[:128], d, 6, "warp").tolist() == [1005, 1032, 1064, 1096, 4, 5]
    # lane 0's third entry moves to the front of its list
    d = np.ones(128, np.uint32)
    d[64] = 50
    assert oracle.select_states_gpu(p[:128], d, 5, "warp").tolist() == [1064, 1000, 1032, 1096, 4]
    # a fifth entry with a larger divergence pushes lane 0's fourth out
    d = np.ones(160, np.uint32)
    d[128] = 9
    assert oracle.select_states_gpu(p, d, 4, "warp").tolist() == [1128, 1000, 1032, 1064]


def test_gpu_insert_selection_is_a_stable_sort_by_divergence():
    rng = np.random.default_rng(5)
    for H, L in [(50, 8), (300, 64), (1000, 17), (12, 20)]:
        p = rng.permutation(H).astype(np.uint32)
        d = rng.integers(0, 6, H).astype(np.uint32)
        order = sorted(range(H), key=lambda i: -int(d[i]))  # Python's sort is stable
        want = [int(p[order[i]]) if i < H else i % H for i in range(L)]
        assert oracle.select_states_gpu(p, d, L, "insert").tolist() == want


@pytest.mark.parametrize("H,L", [(200, 8), (12000, 8), (300, 40)])
def test_state_table_gpu_pairs_and_alleles(H, L):
    ref, prefix, div = _pbwt_rows(H)
    h0, h1, a0, a1 = oracle.state_table_gpu(ref, prefix, div, L)
    for m in range(ref.shape[0]):
        sel = oracle.select_states_gpu(prefix[m], div[m], L)
        assert h0[m].tolist() == sel.tolist()
        assert h1[m].tolist() == [(int(h) + 1) % H for h in sel]
        assert a0[m].tolist() == [ref[m, h] for h in sel]
        assert a1[m].tolist() == [ref[m, (int(h) + 1) % H] for h in sel]


# ---------------------------------------------------------------- transitions
def test_transition_limits():
    # 50 kb at 1 cM/Mb, Ne = 1e4, rho_rate 4: exponent -80, rho == 1.0 in double,
    # so stay == switch == 1/(2 Ne).
    g = oracle.genetic_map(np.array([50000, 100000], np.uint64))
    Td, To = oracle.transitions(g, 10000)
    assert Td[0] == To[0] == np.float32(math.log10(1 / 20000))
    # coincident markers: rho = 0, stay = 1, switch below 1e-300 -> -999
    Td, To = oracle.transitions(np.array([0.5, 0.5]), 10000)
    assert Td[0] == 0.0 and To[0] == -999.0
    # decreasing map clamps the distance to 0
    Td2, To2 = oracle.transitions(np.array([0.5, 0.4]), 10000)
    assert Td2[0] == 0.0 and To2[0] == -999.0


@pytest.mark.parametrize("ne", [1000, 10000, 100000])
def test_transition_probability_identities(ne):
    # stay + (2Ne - 1) switch = 1 and 2Ne switch = 1 - exp(-4 Ne d rho_rate)
    d = np.array([1e-7, 3e-6, 2e-5, 1e-4])
    g = np.concatenate([[0.0], np.cumsum(d)])
    Td, To = oracle.transitions(g, ne)
    stay = 10.0 ** Td.astype(np.float64)
    sw = 10.0 ** To.astype(np.float64)
    assert np.allclose(stay + (2 * ne - 1) * sw, 1.0, rtol=1e-6)
    assert np.allclose(2 * ne * sw, -np.expm1(-4 * ne * d * 4.0), rtol=1e-6)


def test_genetic_map_cm_precedence():
    pos = np.array([1000, 2000, 3000], np.uint64)
    g = oracle.genetic_map(pos, np.array([0.0, 0.5, 0.0]))
    assert g.tolist() == [1000 / 1e8, 0.005, 3000 / 1e8]


# ------------------------------------------------------- forward-backward pins
def lse(xs):
    m = max(xs)
    return m + math.log(sum(math.exp(x - m) for x in xs))


def emission_matrix(a0, a1, gl_s):
    M, L = a0.shape
    E = np.empty((M, L))
    for m in range(M):
        for k in range(L):
            s = int(a0[m, k]) + int(a1[m, k])
            E[m, k] = gl_s[m, 0 if s == 0 else (1 if s == 1 else 2)]
    return E


def brute_posterior(a0, a1, gl_s, Td, To, per_marker, forward_floor=True):
    """Path enumeration.  Forward: every path ending in state k, plus the
    paths the -999 floor starts at each marker r >= 1 (the reference adds the
    floor to the forward vector rescaled to maximum 0, so in unscaled terms a
    path starting at r carries -999 + max_i A_{r-1}[i]); backward: the partial
    paths from k, the ones stopping before the last marker weighted by the
    -999 floor (e^-999), the full ones by 1.  forward_floor=False drops the
    floor-started forward paths (a planted mistake for the tests)."""
    M, L = a0.shape
    E = emission_matrix(a0, a1, gl_s)

    def T(r, i, j):  # step into marker r
        t = (r - 1) if per_marker else 0
        return float(Td[t] if i == j else To[t])

    post = np.empty((M, L))
    fwd_all = []
    for m in range(M):
        lf = []
        for k in range(L):
            terms = []
            for path in itertools.product(range(L), repeat=m):
                s = list(path) + [k]
                w = E[0, s[0]] + sum(T(r, s[r - 1], s[r]) + E[r, s[r]] for r in range(1, m + 1))
                terms.append(w)
            for r0 in range(1, m + 1) if forward_floor else ():
                start = -999.0 + max(fwd_all[r0 - 1])
                for path in itertools.product(range(L), repeat=m - r0):
                    s = list(path) + [k]  # states at markers r0..m
                    w = start + E[r0, s[0]] + sum(T(q, s[q - r0 - 1], s[q - r0]) + E[q, s[q - r0]]
                                                  for q in range(r0 + 1, m + 1))
                    terms.append(w)
            lf.append(lse(terms))
        lb = []
        for k in range(L):
            terms = []
            rest = M - 1 - m
            for ell in range(rest + 1):
                c = 0.0 if ell == rest else -999.0
                for path in itertools.product(range(L), repeat=ell):
                    s = [k] + list(path)
                    w = c + sum(T(m + r, s[r - 1], s[r]) + E[m + r, s[r]] for r in range(1, ell + 1))
                    terms.append(w)
            lb.append(lse(terms))
        x = np.array(lf) + np.array(lb)
        x = np.exp(x - x.max())
        post[m] = x / x.sum()
        fwd_all.append(lf)
    return post, lse(fwd_all[-1])


def random_instance(rng, M, L, with_floor=True):
    a0 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    a1 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    if rng.random() < 0.3:
        a1[rng.integers(M), rng.integers(L)] = 255
    gl = rng.uniform(-3.0, 0.0, (1, M, 3)).astype(np.float32)
    if with_floor:  # hard calls: every present class mismatching at some marker
        gl[0, rng.integers(M)] = [-999.0, -999.0, -999.0]
        gl[0, rng.integers(M)] = [0.0, -999.0, -999.0]
    To = rng.uniform(-8.0, -1.0, M - 1).astype(np.float32)
    Td = np.maximum(To, rng.uniform(-1.0, 0.0, M - 1)).astype(np.float32)
    if rng.random() < 0.3:
        Td = To.copy()
    return a0, a1, gl, Td, To


def coincident_instance(rng, M, L):
    """Hard calls with missing markers over coincident intervals (rho = 0:
    stay 0, switch -999, the oracle's transitions of equal positions), some
    intervals ordinary: no switching across the coincident ones, so the
    -999 floor carries the forward and backward vectors there."""
    a0 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    a1 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    pal = np.array([[0.0, -999.0, -999.0], [-999.0, 0.0, -999.0], [-999.0, -999.0, 0.0],
                    [0.0, 0.0, 0.0]], np.float32)
    gl = pal[rng.integers(0, 4, M)][None]
    step = rng.integers(0, 3, M) * 1000  # repeated positions, at least one pair
    step[1 + int(rng.integers(0, M - 1))] = 0
    pos = np.cumsum(step).astype(np.uint64)
    Td, To = oracle.transitions(oracle.genetic_map(pos), 10000)
    return a0, a1, gl, Td, To


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("per_marker", [False, True])
def test_fb_f64_matches_brute_force_at_coincident_markers(seed, per_marker):
    rng = np.random.default_rng(300 + seed)
    M, L = int(rng.integers(2, 6)), int(rng.integers(2, 4))
    a0, a1, gl, Td, To = coincident_instance(rng, M, L)
    if not per_marker:
        Td[0], To[0] = 0.0, -999.0  # the first interval drives every step
    assert (To == -999.0).any()
    ref, ll = brute_posterior(a0, a1, gl[0], Td, To, per_marker)
    out = oracle.fb_f64(a0, a1, gl, Td, To, per_marker=per_marker)
    assert np.abs(out["prob"][0] - ref).max() < 1e-10
    assert abs(out["loglik"][0] - ll) < 1e-9 * max(1.0, abs(ll))
    fast = oracle.fb_f64_fast(a0, a1, gl, Td, To, per_marker=per_marker)
    assert np.abs(fast["prob"][0] - ref).max() < 1e-10
    assert abs(fast["loglik"][0] - ll) < 1e-9 * max(1.0, abs(ll))


def test_forward_floor_changes_the_answer_at_coincident_markers():
    # The forward floor is what carries a coincident run whose states all
    # mismatch: enumerating without it (a planted mistake) must move the
    # posteriors, so the brute force above pins the oracle's forward floor.
    hit = 0
    for seed in range(40):
        rng = np.random.default_rng(500 + seed)
        a0, a1, gl, Td, To = coincident_instance(rng, 5, 3)
        Td[0], To[0] = 0.0, -999.0
        good, _ = brute_posterior(a0, a1, gl[0], Td, To, False)
        bad, _ = brute_posterior(a0, a1, gl[0], Td, To, False, forward_floor=False)
        out = oracle.fb_f64(a0, a1, gl, Td, To, per_marker=False)
        assert np.abs(out["prob"][0] - good).max() < 1e-10
        if np.isfinite(bad).all() and np.abs(bad - good).max() > 1e-3:
            hit += 1
    assert hit >= 3, hit


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("per_marker", [False, True])
def test_fb_f64_matches_brute_force(seed, per_marker):
    rng = np.random.default_rng(100 + seed)
    M, L = int(rng.integers(1, 6)), int(rng.integers(1, 4))
    a0, a1, gl, Td, To = random_instance(rng, M, L)
    ref, ll = brute_posterior(a0, a1, gl[0], Td, To, per_marker)
    out = oracle.fb_f64(a0, a1, gl, Td, To, per_marker=per_marker)
    assert np.abs(out["prob"][0] - ref).max() < 1e-10
    assert abs(out["loglik"][0] - ll) < 1e-9 * max(1.0, abs(ll))
    fast = oracle.fb_f64_fast(a0, a1, gl, Td, To, per_marker=per_marker)
    assert np.abs(fast["prob"][0] - ref).max() < 1e-10


def test_floor_changes_the_answer():
    # A planted mistake (backward without the -999 floor) must be detectable by
    # the brute force: drop the floor terms and the posteriors move.
    rng = np.random.default_rng(7)
    a0, a1, gl, Td, To = random_instance(rng, 5, 3)
    gl[0, 3] = [-999.0, -999.0, -999.0]
    with_floor, _ = brute_posterior(a0, a1, gl[0], Td, To, True)
    out = oracle.fb_f64(a0, a1, gl, Td, To, per_marker=True)
    assert np.abs(out["prob"][0] - with_floor).max() < 1e-10
    # floor-free enumeration differs at the markers before the all-mismatch one
    M, L = a0.shape
    E = emission_matrix(a0, a1, gl[0])
    lb = np.zeros(L)
    for m in range(M - 2, 2, -1):
        lb = np.array([lse([(Td[m] if i == j else To[m]) + E[m + 1, j] + lb[j] for j in range(L)])
                       for i in range(L)])
    assert np.isfinite(lb).all()


def brute_windowed(a0, a1, gl_s, Td, To, per_marker, W, seed_prev=False):
    """Path enumeration of the windowed model: a step into a window start
    w*W (w > 0) carries the state unchanged with weight 1 (no transition, no
    emission); window w > 0 restarts its forward at the first forward's
    normalized vector at its own last marker; the backward may stop at
    marker e with weight e^-999 unless the step out of e is a carry."""
    M, L = a0.shape
    E = emission_matrix(a0, a1, gl_s)

    def step(r, i, j):  # log weight of the step into marker r
        if r % W == 0:
            return 0.0 if i == j else -math.inf
        t = (r - 1) if per_marker else 0
        return float(Td[t] if i == j else To[t]) + E[r, j]

    def fwd(lo, m, init):  # log sum over paths lo..m ending in each state
        out = []
        for k in range(L):
            terms = []
            for path in itertools.product(range(L), repeat=m - lo):
                z = list(path) + [k]
                terms.append(init[z[0]] + sum(step(lo + r, z[r - 1], z[r]) for r in range(1, m - lo + 1)))
            out.append(lse(terms) if max(terms) > -math.inf else -math.inf)
        return np.array(out)

    def bwd(m, k):
        terms = []
        for e in range(m, M):
            if e < M - 1 and (e + 1) % W == 0:
                continue  # no floor term before a carry
            c = 0.0 if e == M - 1 else -999.0
            for path in itertools.product(range(L), repeat=e - m):
                z = [k] + list(path)
                terms.append(c + sum(step(m + r, z[r - 1], z[r]) for r in range(1, e - m + 1)))
        return lse([t for t in terms if t > -math.inf])

    nw = (M + W - 1) // W
    F = []
    for w in range(nw):
        end = min(M, (w + 1) * W) - 1
        f = fwd(0, end, E[0])
        F.append(f - f.max())
    post = np.empty((M, L))
    for m in range(M):
        w = m // W
        a = fwd(0, m, E[0]) if w == 0 else fwd(w * W, m, F[w - 1] if seed_prev else F[w])
        x = a + np.array([bwd(m, k) for k in range(L)])
        x = np.exp(x - x.max())
        post[m] = x / x.sum()
    return post


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("W", [1, 2, 3])
def test_fb_windowed_matches_brute_force(seed, W):
    rng = np.random.default_rng(300 + seed)
    M, L = int(rng.integers(W + 1, 7)), int(rng.integers(1, 4))
    a0, a1, gl, Td, To = random_instance(rng, M, L)
    per_marker = bool(seed % 2)
    ref = brute_windowed(a0, a1, gl[0], Td, To, per_marker, W)
    out = oracle.fb_f64_windowed(a0, a1, gl, Td, To, W, per_marker=per_marker)
    assert np.abs(out["prob"][0] - ref).max() < 1e-10


def test_fb_windowed_seed_is_the_windows_own_final_vector():
    # The backward re-run of window w starts from window w's own final forward
    # vector (imputer.cpp:1120-1142); seeding it from window w-1's instead
    # (the continuous chain) is a different answer the brute force tells apart.
    rng = np.random.default_rng(77)
    a0, a1, gl, Td, To = random_instance(rng, 6, 3, with_floor=False)
    own = brute_windowed(a0, a1, gl[0], Td, To, True, 3)
    prev = brute_windowed(a0, a1, gl[0], Td, To, True, 3, seed_prev=True)
    out = oracle.fb_f64_windowed(a0, a1, gl, Td, To, 3, per_marker=True)
    assert np.abs(out["prob"][0] - own).max() < 1e-10
    assert np.abs(out["prob"][0] - prev).max() > 1e-3


def test_fb_windowed_closed_forms():
    rng = np.random.default_rng(21)
    B, M, L = 2, 30, 4
    a0 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    a1 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    gl = rng.uniform(-4, 0, (B, M, 3)).astype(np.float32)
    To = rng.uniform(-6, -2, M - 1).astype(np.float32)
    Td = np.maximum(To, rng.uniform(-0.5, 0, M - 1)).astype(np.float32)
    # no windowing when M <= W
    for W in (0, M, M + 5):
        a = oracle.fb_f64_windowed(a0, a1, gl, Td, To, W, per_marker=True)
        b = oracle.fb_f64(a0, a1, gl, Td, To, per_marker=True)
        assert np.array_equal(a["prob"], b["prob"]) and np.array_equal(a["loglik"], b["loglik"])
    # W = 1: every window is one marker, so every forward is marker 0's
    # normalized emission and every backward is empty
    a = oracle.fb_f64_windowed(a0, a1, gl, Td, To, 1, per_marker=True)
    for s in range(B):
        e = emission_matrix(a0, a1, gl[s])[0]
        x = np.exp(e - e.max())
        assert np.abs(a["prob"][s] - x / x.sum()).max() < 1e-12
    # W = M - 1: the last window is its own first marker, carried from M - 2;
    # the first window is the plain model on markers 0..M-2
    a = oracle.fb_f64_windowed(a0, a1, gl, Td, To, M - 1, per_marker=True)
    b = oracle.fb_f64(a0[:-1], a1[:-1], gl[:, :-1], Td[:-1], To[:-1], per_marker=True)
    assert np.abs(a["prob"][:, :-1] - b["prob"]).max() < 1e-12
    assert np.abs(a["prob"][:, -1] - a["prob"][:, -2]).max() < 1e-12
    assert np.abs(a["loglik"] - b["loglik"]).max() < 1e-9


def test_fb_uniform_transitions_closed_form():
    # Td == To (rho = 1, the reference's configuration at 50 kb spacing):
    # backward is the same for every state, so posterior = softmax(E_m).
    rng = np.random.default_rng(3)
    M, L = 40, 8
    a0 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    a1 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    codes = rng.integers(0, 4, (2, M))
    pal = np.array([[0, -999, -999], [-999, 0, -999], [-999, -999, 0], [0, 0, 0]], np.float32)
    gl = pal[codes]
    Td = np.full(M - 1, np.float32(math.log10(1 / 20000)))
    out = oracle.fb_f64(a0, a1, gl, Td, Td.copy())
    for s in range(2):
        E = emission_matrix(a0, a1, gl[s])
        x = np.exp(E - E.max(1, keepdims=True))
        assert np.abs(out["prob"][s] - x / x.sum(1, keepdims=True)).max() < 1e-12


def test_fb_reference_integration_case():
    # test_kernels.cu:274-362: B=1, M=5, L=2, every emission and transition
    # ln(1/2); posteriors must sum to 1 (to 0.01) -- here each is exactly 1/2.
    M, L = 5, 2
    h = np.float32(math.log(0.5))
    gl = np.full((1, M, 3), h, np.float32)
    a0 = np.zeros((M, L), np.uint8)
    out = oracle.fb_f64(a0, a0, gl, np.full(M - 1, h), np.full(M - 1, h), want_logpost=True)
    assert np.abs(np.exp(out["logpost"]).sum(-1) - 1).max() < 0.01
    assert np.abs(out["prob"] - 0.5).max() < 1e-12


def test_logsumexp2_known_answers():
    # test_kernels.cu:48-54, in the oracle's fold: log(0.5)+log(0.5), 0.3/0.7,
    # 0.1/0.9 and (-999, log 1) all combine to 0 within 0.01.  Exercised through
    # the FP32 faithful forward with two states and zero transitions.
    for p, q in [(0.5, 0.5), (0.3, 0.7), (0.1, 0.9)]:
        assert abs(lse([math.log(p), math.log(q)])) < 0.01
    assert abs(lse([-999.0, 0.0])) < 0.01


@pytest.mark.parametrize("per_marker", [False, True])
def test_fb_fast_matches_fold_medium(per_marker):
    rng = np.random.default_rng(11)
    B, M, L = 3, 200, 8
    a0 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    a1 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    pal = np.array([[0, -999, -999], [-999, 0, -999], [-999, -999, 0], [0, 0, 0]], np.float32)
    gl = pal[rng.integers(0, 4, (B, M))]
    To = rng.uniform(-6, -2, M - 1).astype(np.float32)
    Td = np.maximum(To, rng.uniform(-0.5, 0, M - 1)).astype(np.float32)
    a = oracle.fb_f64(a0, a1, gl, Td, To, per_marker=per_marker)
    b = oracle.fb_f64_fast(a0, a1, gl, Td, To, per_marker=per_marker)
    assert np.abs(a["prob"] - b["prob"]).max() < 1e-11
    assert np.abs(a["ds"] - b["ds"]).max() < 1e-11
    assert np.abs(a["loglik"] - b["loglik"]).max() < 1e-8


def test_fb_dosage_definition():
    rng = np.random.default_rng(5)
    a0, a1, gl, Td, To = random_instance(rng, 4, 3)
    out = oracle.fb_f64(a0, a1, gl, Td, To)
    p = out["prob"][0]
    assert np.allclose(out["ap"][0, :, 0], (p * (a0 == 1)).sum(1))
    assert np.allclose(out["ap"][0, :, 1], (p * (a1 == 1)).sum(1))
    assert np.allclose(out["ds"][0], out["ap"][0].sum(1))


def test_fp32_faithful_drift_is_bounded_but_nonzero():
    # Plain FP32 log-space arithmetic with beta resting on the -999 floor drifts
    # measurably from FP64; this quantifies it (it is why the device kernel
    # carries the backward scale in FP64).
    rng = np.random.default_rng(9)
    B, M, L = 2, 600, 8
    a0 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    a1 = rng.integers(0, 2, (M, L)).astype(np.uint8)
    pal = np.array([[0, -999, -999], [-999, 0, -999], [-999, -999, 0], [0, 0, 0]], np.float32)
    codes = np.full((B, M), 3)
    codes[:, ::4] = rng.integers(0, 3, (B, (M + 3) // 4))
    gl = pal[codes]
    To = np.full(M - 1, -4.0, np.float32)
    Td = np.full(M - 1, -0.3, np.float32)
    a = oracle.fb_f64(a0, a1, gl, Td, To)
    f = oracle.fb_f32_faithful(a0, a1, gl, Td, To)
    err = np.abs(a["ds"] - f["ds"]).max()
    assert err < 1e-2


# ------------------------------------------------------------------- sampling
def test_philox_known_answers():
    d = json.load(open(os.path.join(GOLDEN, "philox_kat.json")))
    for src in ("random123", "rocrand"):
        for v in d[src]:
            assert oracle.philox4x32_10(v["ctr"], v["key"]).tolist() == v["out"], (src, v)


def test_sample_uniform_mapping():
    d = json.load(open(os.path.join(GOLDEN, "philox_kat.json")))
    # counter (m >> 2, s, stream, 0) -> word m & 3 -> (x >> 8) / 2^24
    for m in range(8):
        for s in (0, 5):
            w = oracle.philox4x32_10([m >> 2, s, 0, 0], [12345, 0])[m & 3]
            assert oracle.sample_uniform(12345, s, m) == np.float32((int(w) >> 8) / 16777216.0)
    assert d["random123"]


def test_sample_from_probs_closed_form():
    p = np.full(4, 0.25, np.float32)
    assert oracle.sample_from_probs(p, 0.0) == 0
    assert oracle.sample_from_probs(p, 0.1) == 0
    assert oracle.sample_from_probs(p, 0.25) == 1   # u < cumsum is strict
    assert oracle.sample_from_probs(p, 0.74) == 2
    assert oracle.sample_from_probs(p, 0.99) == 3
    q = np.array([0.5, np.float32(0.5) - np.float32(2.0 ** -24)], np.float32)
    top = float(np.cumsum(q, dtype=np.float32)[-1])  # FP32 cumsum ends below 1
    assert top < 1.0
    assert oracle.sample_from_probs(q, np.nextafter(np.float32(top), np.float32(1))) == 0
    assert oracle.argmax(np.array([0.1, 0.4, 0.4, 0.1], np.float32)) == 1


# ------------------------------------------------------------ CPU baseline
@pytest.mark.parametrize("seed,L", [(3, 4), (4, 8)])
def test_cpu_reference_impute_is_the_reference_loop(seed, L):
    # The timed CPU baseline (PBWT sweep, per-(sample, marker) std::sort
    # selection, L^2 FP32 fold with sqrt(M) checkpoint replay, inverse-CDF
    # draw) computes what the reference computes: its haplotypes equal the
    # draw replayed over the FP32 reference-order forward-backward (which keeps
    # every forward vector) on the state table the PBWT selects.
    rng = np.random.default_rng(seed)
    M, H, B = 60, 40, 5
    ref = (rng.random((M, H)) < 0.4).astype(np.uint8)
    gl = np.zeros((B, M, 3), np.float32)
    typed = rng.random((B, M)) < 0.3
    call = rng.integers(0, 3, (B, M))
    for g in range(3):
        sel = typed & (call == g)
        gl[sel] = -999.0
        gl[sel, g] = 0.0
    pos = (np.arange(M, dtype=np.uint64) + 1) * 7000
    Td, To = oracle.transitions(oracle.genetic_map(pos), 10000)
    hap = oracle.cpu_reference_impute(ref, gl, Td, To, L, 99, nthreads=2)
    p, _ = oracle.pbwt_build(ref)
    _, _, a0, a1 = oracle.state_table(ref, p, L)
    f = oracle.fb_f32_faithful(a0, a1, gl, Td, To, per_marker=False)
    want = oracle.sample_replay(f["prob"], a0, a1, 99)
    assert np.array_equal(hap, want)
    # and the draw is not trivially constant: the posteriors spread over states
    assert (f["prob"].max(axis=2) < 0.999).any()
