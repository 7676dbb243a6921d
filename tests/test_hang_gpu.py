"""The hang bound (liblizard/core/wg/wg.h:48-55): a handle's wait for its own GPU work longer than the bound takes the
work as hung, fails with LIZ_E_STATE and leaves the handle refusing further calls, as the reference drops a lost
device. The condition is forced from the host: the bound set below what a real decode takes (the work itself ends
normally), then restored; a new handle decodes as before."""
import numpy as np
import pytest

import lizard_amd
from _libs import RATES
from ldpc_cases import batch

LIZ_E_STATE = -7


@pytest.mark.gpu
def test_wait_past_the_bound_is_taken_as_hung():
    rates = [RATES[i % 4] for i in range(2048)]
    rt, llr, offs, _ = batch(rates, [2.0] * 2048, seed=77)
    ref = lizard_amd.LdpcDecoder()
    try:
        want = ref.decode(rt, llr, offs, max_iter=20, stall_it=0)
    finally:
        ref.close()
    dec = lizard_amd.LdpcDecoder()
    try:
        lizard_amd.hang_bound(1e-4)  # 0.1 us: no decode of 2048 codewords ends within it
        with pytest.raises(lizard_amd.LizError) as e:
            dec.decode(rt, llr, offs, max_iter=20, stall_it=0)
        assert e.value.code == LIZ_E_STATE and "hung" in str(e.value)
        lizard_amd.hang_bound(4000)
        with pytest.raises(lizard_amd.LizError) as e:  # the handle stays lost, whatever the bound now
            dec.decode(rt, llr, offs, max_iter=20, stall_it=0)
        assert e.value.code == LIZ_E_STATE
    finally:
        lizard_amd.hang_bound(4000)
        dec.close()
    fresh = lizard_amd.LdpcDecoder()
    try:
        got = fresh.decode(rt, llr, offs, max_iter=20, stall_it=0)
        assert np.array_equal(got["its"], want["its"]) and np.array_equal(got["bits"], want["bits"])
    finally:
        fresh.close()


def test_bound_refused_out_of_range():
    with pytest.raises(lizard_amd.LizError):
        lizard_amd.hang_bound(-1)
