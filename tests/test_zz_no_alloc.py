"""tessera_sv_anneal_static (no device allocation: every buffer in the
library's static per-GPU scratch) against tessera_sv_anneal on the same inputs,
through the C ABI: identical kernels and launch shapes, so results must be
bit-identical. Two calls in a row, the second with a smaller input, on both
code paths of the anneal (n <= 13: state resident in one workgroup's LDS;
n >= 14: streaming passes with per-block partials), so the second call runs on
scratch the first one left behind."""
import numpy as np
import pytest

from tessera_amd import instances as I
from tessera_amd import qa

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert np.array_equal(a.spins, b.spins)
    assert a.energy == b.energy
    assert a.expected_energy == b.expected_energy
    assert a.success_prob == b.success_prob
    assert np.array_equal(a.entropy_trace, b.entropy_trace)
    assert np.array_equal(a.energy_trace, b.energy_trace)


@pytest.mark.parametrize("sizes", [((16, 40), (14, 25)), ((12, 60), (7, 30)), ((16, 30), (9, 20))])
def test_static_scratch_matches_allocating_entry(sizes):
    for n, steps in sizes:
        model = I.random_regular3(n, 500 + n) if n % 2 == 0 else I.chain_nnn(n, 500 + n)
        sched = I.linear(steps)
        print(f"static vs allocating: n={n} steps={steps}", flush=True)
        ref = qa.anneal_statevector(model, sched, 0.1)
        got = qa.anneal_statevector_static(model, sched, 0.1)
        _same(got, ref)


def test_static_scratch_rejects_oversized_state():
    model = I.chain_nnn(17, 3)
    with pytest.raises(qa.TesseraError) as e:
        qa.anneal_statevector_static(model, I.linear(4), 0.1)
    assert e.value.status == -2
