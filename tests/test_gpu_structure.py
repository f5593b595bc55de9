"""GPU cell structure and weight state against the oracle: bit-exact integers, FP64 sums."""
import numpy as np
import pytest

from gpu_util import announce
from oracle.oracle import OracleSolver
from test_oracle import designs

pytestmark = pytest.mark.gpu


def _compare(codes, gs, lim, w=None):
    import fereg_amd as F
    o = OracleSolver(codes, gs, weights=w, small_limit=lim)
    a = F.Absorber(codes, gs, weights=w, small_limit=lim)
    oi, ai = o.info, a.info
    for f in ("n", "g1", "ncells", "nb", "m", "m_small", "small_work", "fe1", "norm_fe1", "has_cell_obs_off"):
        assert getattr(oi, f) == getattr(ai, f), f
    for x, y in zip(o.blocks(), a.blocks()):
        np.testing.assert_array_equal(x, y)
    so, sa = o.structure(), a.structure()
    for k in ("pi", "seg_off", "cell_codes"):
        np.testing.assert_array_equal(so[k], sa[k], err_msg=k)
    if so["cell_obs_off"] is None:
        assert sa["cell_obs_off"] is None
    else:
        np.testing.assert_array_equal(so["cell_obs_off"], sa["cell_obs_off"])
    np.testing.assert_array_equal(a.row_order(), so["pi"])
    for b in range(oi.nb):
        po, fo = o.level_cells(b)
        pa, fa = a.level_cells(b)
        np.testing.assert_array_equal(po, pa)
        np.testing.assert_array_equal(fo, fa)
    wo, wa = o.weight_state(), a.weight_state()
    np.testing.assert_allclose(wa["n1"], wo["n1"], rtol=1e-13)
    np.testing.assert_allclose(wa["nr"], wo["nr"], rtol=1e-13)
    np.testing.assert_allclose(wa["diag"], wo["diag"], rtol=1e-9, atol=1e-9 * max(1.0, wo["nr"].max(initial=1.0)))
    np.testing.assert_array_equal(wa["redundant"], wo["redundant"])
    assert abs(wa["sw"] - wo["sw"]) <= 1e-12 * wo["sw"]
    return o, a


@pytest.mark.parametrize("di", range(8))
def test_structure_bit_exact(di):
    codes, gs, limits = designs()[di]
    for lim in limits:
        announce("structure design", di, "small_limit", lim)
        _compare(codes, gs, lim)


@pytest.mark.parametrize("sort", [False, True])
def test_structure_large_random(sort):
    rng = np.random.default_rng(31)
    n = 200_000
    gs = [20_000, 10, 5000]
    codes = [rng.integers(0, g, n).astype(np.uint32) for g in gs]
    if sort:
        idx = np.lexsort(codes[::-1])
        codes = [c[idx] for c in codes]
    announce("structure large", n, "sorted", sort)
    _compare(codes, gs, 4096)


@pytest.mark.parametrize("maxlen,gap", [(16, 0), (32, 0), (64, 0), (65, 0), (16, 3), (16, 100)])
def test_structure_ascending_fe1_levels(maxlen, gap):
    """FE1 ascending in kept order: levels of up to 64 rows are sorted in registers by their block
    codes (ties broken by row, as the stable radix sort does); a 65-row level, or a run of more
    than 64 empty FE1 levels, takes the radix sort."""
    rng = np.random.default_rng(maxlen + gap)
    lens = rng.integers(1, maxlen + 1, 3000)
    lens[7] = maxlen
    fe1 = np.repeat(np.arange(len(lens)), lens)
    fe1 = np.where(fe1 >= 500, fe1 + gap, fe1).astype(np.uint32)  # `gap` empty levels after level 499
    g1 = len(lens) + gap
    n = len(fe1)
    firm = rng.integers(0, 50, n).astype(np.uint32)
    year = rng.integers(0, 4, n).astype(np.uint32)
    announce("structure ascending fe1", maxlen, gap)
    _compare([fe1, firm, year], [g1, 50, 4], 4096)


def test_structure_panel_and_repeated_cells():
    rng = np.random.default_rng(5)
    i = np.arange(50_000)
    indiv, year = (i // 10).astype(np.uint32), (i % 10).astype(np.uint32)
    firm = rng.integers(0, 700, len(i)).astype(np.uint32)
    announce("structure panel")
    _compare([indiv, firm, year], [5000, 700, 10], 4096)
    # repeated cells: duplicated rows, weights
    codes = [np.repeat(indiv[:20000], 2), np.repeat(firm[:20000], 2), np.repeat(year[:20000], 2)]
    w = rng.uniform(0.5, 2.0, len(codes[0]))
    announce("structure repeated cells weighted")
    _compare(codes, [2000, 700, 10], 4096, w)


def test_wide_keys_use_several_words():
    """Codes wider than 64 bits in total (LSD over 64-bit words)."""
    rng = np.random.default_rng(8)
    n = 6000
    codes = [rng.integers(0, 900, n).astype(np.uint32)]
    gs = [1 << 20]
    for _ in range(4):
        codes.append((rng.integers(0, 6, n) << 14).astype(np.uint32))
        gs.append(1 << 18)
    announce("structure wide keys")
    o, a = _compare(codes, gs, 0)
    # level records too wide to ride in the sort keys: the gathered fallback feeds the operator
    for r in (1, 2):
        p = rng.standard_normal(o.info.m * r)
        yo = o.apply(p, r)
        announce("wide keys apply", r)
        np.testing.assert_allclose(a.apply(p, r), yo, rtol=0, atol=1e-12 * np.abs(yo).max())


def test_invalid_inputs_are_rejected():
    import fereg_amd as F
    with pytest.raises(F.FeregError) as e:
        F.Absorber([np.array([0, 5], np.uint32)], [3])
    assert e.value.rc == 198
    with pytest.raises(F.FeregError) as e:
        F.Absorber([np.array([0, 1], np.uint32)], [2], weights=np.array([1.0, -1.0]))
    assert e.value.rc == 198
