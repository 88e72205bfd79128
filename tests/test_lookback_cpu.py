"""CPU model of the index kernel's carry-in protocol (tests/harness/
lookback_model.cpp): threads as the persistent grid's blocks claim tiles from
per-class tickets (block b takes the tiles of class b % npart, npart the
largest power of two up to 8 and the grid), publish aggregates, walk back over
tiles and 64-tile groups, and
publish inclusive prefixes, with random delays at every step.  Every launch
must end and every carry-in must equal the sequential prefix: for one launch
over all tiles, and for the staging pipeline's form -- consecutive launches
over [tile_begin, end) with the tickets reset before each, tile states kept,
grids smaller than the launch, and speculating tiles that publish their prefix
only after their own walk."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("harness") / "liblookback.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread",
                    os.path.join(ROOT, "tests", "harness", "lookback_model.cpp"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    L.lb_set_parts.argtypes = [ctypes.c_int]
    L.lb_set_claim.argtypes = [ctypes.c_int]
    L.lb_run.restype = ctypes.c_int
    L.lb_run.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                         ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
    return L


def run(lib, ntiles, chunk, grid, spec_frac, deferred, seed, delay_pct=2, timeout_ms=30000):
    rng = np.random.default_rng(seed)
    spec = (rng.random(ntiles) < spec_frac).astype(np.uint8)
    bounds = np.array(list(range(0, ntiles, chunk)) + [ntiles], dtype=np.int64)
    return lib.lb_run(ntiles, bounds.ctypes.data, len(bounds) - 1, grid, spec.ctypes.data, int(deferred), delay_pct,
                      seed, timeout_ms)


@pytest.mark.parametrize("spec_frac,deferred", [(0.0, False), (0.5, False), (0.9, True), (0.97, False)])
def test_one_launch(lib, spec_frac, deferred):
    for seed in range(2):
        assert run(lib, 4000, 4000, 32, spec_frac, deferred, seed) == 0, (spec_frac, deferred, seed)


@pytest.mark.parametrize("chunk", [1365, 21, 63, 65, 1])
def test_staging_pipeline_launches(lib, chunk):
    """1 GiB in 48 KiB tiles is 21 845 tiles; 64 MiB chunks give 1365 tiles a
    launch (the case that did not finish on the GPU), 1 MiB chunks 21 (the
    cases that matched).  Chunks aligned and not aligned to the 64-tile
    groups, and one-tile launches."""
    ntiles = 21845 if chunk == 1365 else 2000
    for spec_frac in (0.97, 0.0):
        assert run(lib, ntiles, chunk, 32, spec_frac, False, 11 + chunk) == 0, (chunk, spec_frac)


def test_grid_of_one(lib):
    """A single block claims every tile in order: each walk only reads tiles
    that are already done."""
    assert run(lib, 3000, 700, 1, 0.5, False, 5, delay_pct=0) == 0


@pytest.mark.parametrize("grid", [2, 3, 5, 8, 9, 13])
def test_ticket_classes(lib, grid):
    """The kernel claims from one ticket (kTicketParts = 1).  With up to 8
    classes, grids below 8 blocks split the tiles into 2 or 4 classes (every
    class has a block) and grids of 8 or more into 8; with every block running,
    as here, every launch ends."""
    try:
        for parts in (1, 8):
            lib.lb_set_parts(parts)
            for spec_frac in (0.0, 0.6):
                assert run(lib, 1500, 1500, grid, spec_frac, False, 31 + grid) == 0, (grid, parts, spec_frac)
    finally:
        lib.lb_set_parts(1)


@pytest.mark.parametrize("grid", [1, 3, 32])
def test_two_tiles_per_claim(lib, grid):
    """The 2-level variant claims two consecutive tiles per atomic: one launch
    and the staging pipeline's consecutive launches, speculating and not."""
    try:
        lib.lb_set_claim(2)
        for spec_frac in (0.0, 0.7, 0.97):
            assert run(lib, 3000, 3000, grid, spec_frac, False, 41 + grid) == 0, (grid, spec_frac)
            assert run(lib, 3000, 63, grid, spec_frac, False, 43 + grid) == 0, (grid, spec_frac)
    finally:
        lib.lb_set_claim(1)
