"""CPU-side checks of the product: the C ABI library loads and exports every entry point
declared in include/gpvmc.h, and the host-side model setup matches the oracle."""
import os
import re

import numpy as np
import pytest

import oracle_ref as O
from gpvmc_amd import binding, model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "gpvmc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(gpvmc_\w+)\s*\(", src, re.M)))


def test_header_matches_binding():
    names = declared_entry_points()
    assert len(names) >= 15
    assert set(names) == set(binding.SIGNATURES)


def test_library_loads_and_exports_every_symbol():
    lib = binding.load_library()
    for name in declared_entry_points():
        assert hasattr(lib, name), name
    assert lib.gpvmc_version() == 1


def test_create_rejects_bad_config_without_gpu():
    # argument checks run before any device call
    with pytest.raises(binding.GpvmcError, match="L must be"):
        binding.Batch(3, 4)
    with pytest.raises(binding.GpvmcError, match="n_walkers"):
        binding.Batch(8, 0)
    with pytest.raises(binding.GpvmcError, match="no sweep kernel"):
        binding.Batch(14, 4)  # N = 98 has no launch shape


@pytest.mark.parametrize("L,phi,neel,ne,bc", [(4, 0.25, 0.1, 0.0, (1, 1)), (8, 0.1, 0.3, 0.0, (1, 1)),
                                              (6, 0.3, 0.05, 1.5, (0, 1)), (16, 0.2333, 0.0, 0.0, (1, 1))])
def test_orbital_table_matches_oracle(L, phi, neel, ne, bc):
    a = model.orbital_table(L, phi * np.pi, neel, ne, *bc)
    b = O.orbitals(L, phi * np.pi, neel, ne, *bc)
    np.testing.assert_allclose(a, b, atol=1e-13, rtol=0)
    np.testing.assert_allclose(model.ground_occ(a), O.ground_occ(L, b), atol=1e-13, rtol=0)


@pytest.mark.parametrize("L", [4, 6, 8, 16])
def test_lattice_tables_match_oracle(L):
    assert np.array_equal(model.edges(L), O.edges(L))
    assert np.array_equal(model.neighbours(L), O.neighbours(L))


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32_10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        got = model.philox4x32_10(np.array([ctr], np.uint64), np.array([key], np.uint32))[0]
        assert tuple(int(v) for v in got) == want


def test_philox_uniforms_open_interval_and_counter_layout():
    u = model.philox_uniforms(12345, [0, 7], 10, 50)
    assert u.shape == (2, 50, 4)
    assert (u > 0).all() and (u < 1).all()
    # counter (step, walker): a later window continues the same stream
    v = model.philox_uniforms(12345, [7], 30, 10)
    np.testing.assert_array_equal(u[1, 20:30], v[0])


def test_random_spins_half_filled():
    s = model.random_spins(8, 5, 3)
    assert s.shape == (5, 64)
    assert (s.sum(axis=1) == 32).all()
    assert len({tuple(x) for x in s}) == 5
