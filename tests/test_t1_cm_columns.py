"""The context-modelling kernel solves each coding pass for a whole column of a code-block at
once (libgail_amd/csrc/t1_cm_masks.h): which rows it codes, which turn significant, and every
coded row's neighbourhood as the scan sees it. tests/native/t1cm_emu.cpp runs that
formulation on the host, one lane per column, and compares each pass's symbols with a plain
sequential scan of ISO/IEC 15444-1 D.3 (the order t1_ref.cpp:52-286 codes in). The GPU tests
then compare the kernel's codewords with the oracle's."""
import ctypes as C
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "native", "libt1cm_emu.so")


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        pytest.skip("tests/native/libt1cm_emu.so not built (make -C tests/native)")
    L = C.CDLL(EMU)
    L.t1cm_check.restype = C.c_int
    L.t1cm_check.argtypes = [C.c_uint32, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_longlong), C.c_int]
    L.t1cm_zc_table_check.restype = C.c_int
    return L


def test_zero_coding_contexts_match_table_d1(emu):
    """The bit-sliced zero-coding context of every neighbourhood (256) in all four
    orientations equals Table D.1."""
    assert emu.t1cm_zc_table_check() == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_whole_column_passes_match_the_sequential_scan(emu, seed):
    """Random blocks (sizes 1..32, all orientations, sparse to dense magnitudes): every
    pass's symbol bytes from the whole-column formulation equal the sequential scan's."""
    msg = C.create_string_buffer(512)
    stats = (C.c_longlong * 3)()
    bad = emu.t1cm_check(seed, 3000, msg, 512, stats, 0)
    symbols, passes, rounds = list(stats)
    print(f"[cm] seed {seed}: {passes} passes, {symbols} symbols, {rounds} fixed-point rounds")
    assert symbols > 1_000_000 and passes > 10_000
    assert bad == 0, msg.value.decode()


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_mistakes_are_caught(emu, plant):
    """A below-left neighbour seen across a stripe boundary, a fixed point stopped after one
    round, or the signs above and below swapped, each planted in the host replica, makes
    blocks differ from the sequential scan."""
    msg = C.create_string_buffer(512)
    stats = (C.c_longlong * 3)()
    assert emu.t1cm_check(7, 500, msg, 512, stats, plant) > 0
