"""The device side of the numerics contract that only the GPU can show."""
import pytest

import busseq_gpu_amd as bq

pytestmark = pytest.mark.gpu


def test_device_reciprocal_equals_the_ieee_quotient_on_every_normal_float():
    # bq_rcpf (v_rcp_f32 plus one Newton step of z held to [2^-126, 2^125]) must
    # return the host's 1.0f / z of the held value for every normal z below 2^126;
    # the single-precision terms of every pass take their reciprocals from it
    assert bq.rcp_check(0) == 0
