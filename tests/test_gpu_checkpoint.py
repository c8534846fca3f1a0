"""KITPCHK1 checkpoints on the GPU, through the level C-ABI: the device
payload equals the oracle's bit for bit (36 components, comp_base 25 as in
the reference's example), the written file is byte-identical to one written
from the oracle payload, and a checkpoint restores into another box layout
(Main_GRHDLevelDataBinaryPNKITPDisk.cpp:1385-1586; README: the rank count
may differ from the run that wrote it)."""
import numpy as np
import pytest
import torch

import oracle
from grdzhadzha_amd.checkpoint import (pack_level, read_checkpoint_state,
                                       write_checkpoint,
                                       write_checkpoint_file)
from grdzhadzha_amd.level import LevelOperator
from grdzhadzha_amd.options import CircularBinaryPNParams
from test_checkpoint_cpu import (CFG, fabs_from_field, field_value,
                                 restart_record, tiles)

pytestmark = pytest.mark.gpu

BG = CircularBinaryPNParams(mass_1=0.5, mass_2=0.5, separation=8.0)


def level(boxes, ctx):
    return LevelOperator(BG, boxes, ncomp=36, comp_base=25, context=ctx)


def test_device_payload_equals_oracle(hip_context):
    rng = np.random.default_rng(7)
    boxes = tiles(CFG, 3, 2)
    fabs = [rng.standard_normal((36,) + b.storage_shape) for b in boxes]
    lev = level(boxes, hip_context)
    got = pack_level(lev, lev.pack(fabs), CFG).cpu().numpy()
    lo, hi = CFG.domain()
    want = oracle.checkpoint_pack(boxes, fabs, lo, hi, CFG.num_ghosts)
    assert np.array_equal(got, want)
    # tiles that leave cells uncovered: zeros there, as the reference's
    # zero-initialised, summed buffer
    part = boxes[:4]
    got = pack_level(level(part, hip_context),
                     level(part, hip_context).pack(fabs[:4]), CFG)
    want = oracle.checkpoint_pack(part, fabs[:4], lo, hi, CFG.num_ghosts)
    assert np.array_equal(got.cpu().numpy(), want)


def test_file_and_restore_into_another_layout(hip_context, tmp_path):
    fld = field_value(CFG, 36)
    boxes = tiles(CFG, 2, 2)
    lev = level(boxes, hip_context)
    pool = lev.pack(fabs_from_field(CFG, boxes, fld))
    path = str(tmp_path / "gpu.bin")
    write_checkpoint(path, lev, pool, CFG, restart_record())
    ref = str(tmp_path / "oracle.bin")
    write_checkpoint_file(ref, CFG, restart_record(), fld.ravel(), 36)
    assert open(path, "rb").read() == open(ref, "rb").read()

    other = tiles(CFG, 1, 3)
    lev2 = level(other, hip_context)
    pool2 = torch.full((lev2.pool_size,), -1.0, dtype=torch.float64,
                       device="cuda")
    r = read_checkpoint_state(path, lev2, pool2, CFG)
    assert r == restart_record()
    for got, want in zip(lev2.unpack(pool2),
                         fabs_from_field(CFG, other, fld)):
        assert np.array_equal(got, want)
