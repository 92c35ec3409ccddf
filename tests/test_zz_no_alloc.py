"""fdmamr_max_wavespeed_noalloc (the entry point that allocates no device memory) against
fdmamr_max_wavespeed on the same state: bit-identical maxima (a maximum does not depend on
the reduction order), on two calls in a row with the second on a smaller mesh, then the
first mesh again (the shared static scratch keeps the last call's partials)."""
import pytest

import fdmamr as F
from helpers import riemann_c2

pytestmark = pytest.mark.gpu


def test_noalloc_wavespeed_matches_allocating_entry_point():
    big = F.Solver(riemann_c2(h=1.0 / 128, lvl_amr=2, pre_adapt=2)).setup()
    small = F.Solver(F.Config()).setup()
    assert big.n_leaf > 20000 and small.n_leaf == 400
    big.run_steps(2)
    for s in (big, small, big):
        want = s.max_wavespeed()
        got = s.max_wavespeed_noalloc()
        print(s.n_leaf, want, got)
        assert got == want
        assert got[0] > 0.0 and got[1] > 0.0
    big.close()
    small.close()
