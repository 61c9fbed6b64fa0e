"""The bandwidth reference kernels behind bench.py's roofline figures (lee_copy_bench,
lee_stream_bench): they run, report a positive time per launch, and refuse what they do not
support instead of launching it."""
import pytest

pytestmark = pytest.mark.gpu


def test_reference_kernels_time_and_refuse(lee):
    d = lee.deck.box(20, 20, 16, 0.05, time_domain=4)
    s = lee.Solver(d, precision=32)
    L = lee.lib()
    print("[gpu] reference kernels", flush=True)
    assert L.lee_copy_bench(s.h, 1 << 20, 3) > 0
    for reads, writes in [(1, 1), (4, 4), (8, 4), (12, 8), (20, 8), (28, 8)]:
        ms = L.lee_stream_bench(s.h, 1 << 18, reads, writes, 3)
        assert ms > 0, (reads, writes)
    # unsupported stream counts and sizes: an error value, nothing launched
    for reads, writes in [(2, 1), (0, 1), (29, 8), (12, 9)]:
        assert L.lee_stream_bench(s.h, 1 << 18, reads, writes, 3) == -1.0, (reads, writes)
    assert L.lee_stream_bench(s.h, 0, 12, 8, 3) == -1.0
    assert L.lee_stream_bench(s.h, 1 << 18, 12, 8, 0) == -1.0
    # the solver still steps after them
    s.frame_setup(0)
    s.step(2)
    s.close()
