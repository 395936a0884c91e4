"""Host-side checks of the drop-in modules (CPU only): the block width each
reference-named edge-map op assumes, and the part-span validation that
refuses a structure built with another width."""
import inspect

import pytest
import torch

import att_amd.compat as compat
from att_amd import ops
from oracle import block as B
from tests.gpu_util import TOY_COL, TOY_ROW


def test_default_width_by_precision():
    """fp16 ops read 8-wide maps, tf32 ops 4-wide ones (the widths the
    reference's datasets build with), when no width is given."""
    assert compat.DEFAULT_WIDE == {"fp16": 8, "tf32": 4}
    for ver in (1, 3, 4):
        for mod in (compat.MODULES[f"ATT_SpMM_v{ver}"], compat.MODULES[f"ATT_SDDMM_v{ver}"]):
            for name, fn in vars(mod).items():
                if not callable(fn) or name.startswith("_"):
                    continue
                assert inspect.signature(fn).parameters["wide"].default is None, name
    assert compat._wide_of("fp16", None) == 8 and compat._wide_of("tf32", None) == 4
    assert compat._wide_of("tf32", 16) == 16


@pytest.mark.parametrize("wide", [4, 8, 16])
def test_part_spans_of_reference_builds(wide):
    """Every structure the reference's layout produces passes the check at its
    own width and at the divisors of it; the toy graph's 12-column window fails
    at a larger width."""
    rp, *_ = B.block_balance(TOY_ROW, TOY_COL, None, 8, wide, 2, B.MAP)
    rp = torch.as_tensor(rp, dtype=torch.int32)
    for w in (d for d in (4, 8, 16) if d <= wide and wide % d == 0):
        ops.check_part_spans(rp, w)
    if wide == 4:
        with pytest.raises(ValueError, match="wide=8"):
            ops.check_part_spans(rp, 8)


def test_part_span_check_is_cached_until_written():
    rp = torch.tensor([0, 8, 8, 16, 16], dtype=torch.int32)
    ops.check_part_spans(rp, 8)
    assert rp._attgnn_spans_ok[0] == 8
    rp[2] = 12                     # in-place write bumps the version counter: checked again
    rp[3] = 12
    rp[4] = 12
    with pytest.raises(ValueError):
        ops.check_part_spans(rp, 8)
    with pytest.raises(ValueError, match="part 1 spans 4"):
        ops.check_part_spans(torch.tensor([0, 8, 8, 12, 12], dtype=torch.int32), 8)
