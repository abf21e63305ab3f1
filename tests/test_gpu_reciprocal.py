"""The triangle test's reciprocal of its determinant (recip_rn: the hardware reciprocal and one fused correction,
for 2^-14 < |det| < 2^126, the range every determinant the test divides by lies in, as it requires |det| >
1e-4) is the correctly rounded 1.0f / det of the reference (GPURendererKernel.java:1569) for every float in that
range, both signs: 2.35 billion values compared bit for bit on the GPU (df_check_reciprocal)."""

import pytest

import dayflower_amd as dfa

pytestmark = pytest.mark.gpu


def test_reciprocal_is_the_division_everywhere_it_is_used(gpu):
    tested, bad, first = dfa.check_reciprocal()
    assert tested == 2 * (0x7E800000 - 0x38800001)
    assert bad == 0, (bad, hex(first))
