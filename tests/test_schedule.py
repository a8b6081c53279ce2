"""The multires schedule of reconstruct() (src/MBIR_4D/XT_Engine.c:111-141,
one node): number of stages from the column count, mult_xy doubling while it
divides the columns, mult_z doubling while rows / (2 mult_z) >= 2.  The
schedule comes from the library's tb_multires_schedule, the code
reconstruct() runs (no GPU needed).

Known answers are the schedules the reference yields for the workloads'
shapes; the C2 one is also what reconstruct() runs in the GPU tests."""
import pytest

from timbir_amd.batch import multires_schedule


@pytest.mark.parametrize("rows,cols,expect", [
    (4, 128, [(2, 2), (1, 1)]),                          # plain 3D, 128 columns
    (8, 256, [(4, 4), (2, 2), (1, 1)]),                  # 4D interlaced, 256 columns
    (32, 512, [(8, 8), (4, 4), (2, 2), (1, 1)]),         # 512 columns, 32 rows
    (4, 512, [(8, 2), (4, 2), (2, 2), (1, 1)]),          # mult_z stops at rows / 2
    (4, 192, [(2, 2), (1, 1)]),                          # log2(192 / 64) + 1 = 2.58 -> 2 stages
    (4, 96, [(2, 2), (1, 1)]),                           # at least 2 stages
    (6, 256, [(4, 2), (2, 2), (1, 1)]),                  # 6 rows: mult_z 2 (6 % 4 != 0)
    (16, 16384, [(128, 8), (64, 8), (32, 8), (16, 8), (8, 8), (4, 4), (2, 2), (1, 1)]),  # 8 max
    (4, 384, [(4, 2), (2, 2), (1, 1)]),                  # log2(384 / 64) + 1 = 3.58 -> 3 stages
])
def test_schedule_known_answers(rows, cols, expect):
    assert multires_schedule(rows, cols) == expect


def test_schedule_invariants():
    for rows in (4, 6, 8, 16, 32):
        for cols in (64, 96, 128, 192, 256, 384, 512, 1024, 2048):
            sched = multires_schedule(rows, cols)
            assert 2 <= len(sched) <= 8
            assert sched[-1] == (1, 1)
            for (mx, mz), (nx, nz) in zip(sched, sched[1:]):
                assert mx == 2 * nx and mz in (nz, 2 * nz)
            for mx, mz in sched:
                assert cols % mx == 0 and rows % mz == 0
