"""Bounds and alignment of the 24-bit row layout (no GPU).

A row of ld entries (ld a multiple of 64) takes 3 ld bytes, cut into blocks
of up to 512 entries; a block of nb entries holds its entries' high 16 bits
(2 nb bytes) and then their low 8 bits (nb bytes).  The GEMV reads entries
j .. j+7 (j a multiple of 8) with one 16-byte load of high parts and one
8-byte load of low parts; the set-up store writes each entry's 2 + 1 bytes.
This restates f24_offsets (nsim_dev_common.h) and checks, for every row
length the operators use, that every load stays inside its row and block and
is aligned to its width (rows start at multiples of 3 ld bytes), and that the
entries' bytes tile the row exactly once."""
import re
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOCK = 512


def offsets(j, ld):
    j0 = j & ~(BLOCK - 1)
    nb = min(BLOCK, ld - j0)
    return 3 * j0 + 2 * (j - j0), 3 * j0 + 2 * nb + (j - j0), j0, nb


def test_block_size_matches_the_device_header():
    txt = open(os.path.join(ROOT, "nmag-src_amd", "csrc", "nsim_dev_common.h")).read()
    assert re.search(r"constexpr int kF24Block = %d;" % BLOCK, txt)
    # the formula restated here
    assert "*hi = 3 * (size_t)j0 + 2 * (size_t)(j - j0);" in txt
    assert "*lo = 3 * (size_t)j0 + 2 * (size_t)nb + (size_t)(j - j0);" in txt


@pytest.mark.parametrize("ld", [64, 128, 448, 512, 576, 1024, 1088, 7616, 19968, 20480])
def test_loads_and_stores_stay_in_row_and_block(ld):
    row_bytes = 3 * ld
    assert (row_bytes % 16) == 0  # every row starts 16-byte aligned
    cover = np.zeros(row_bytes, dtype=np.int32)
    for j in range(ld):
        hi, lo, j0, nb = offsets(j, ld)
        cover[hi:hi + 2] += 1
        cover[lo] += 1
        assert 3 * j0 <= hi and hi + 2 <= 3 * j0 + 2 * nb
        assert 3 * j0 + 2 * nb <= lo < 3 * (j0 + nb) <= row_bytes
    assert np.all(cover == 1)
    for j in range(0, ld, 8):  # the GEMV's eight-entry loads
        hi, lo, j0, nb = offsets(j, ld)
        assert hi % 16 == 0 and lo % 8 == 0
        assert hi + 16 <= 3 * j0 + 2 * nb
        assert lo + 8 <= 3 * (j0 + nb) <= row_bytes
