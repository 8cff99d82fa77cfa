"""GPU tests of the schedule autotuner: the tuned schedule is never slower than the
planner's default in the tuner's own timing, a plan created with SPLIT_TUNED uses it, and
its results keep the precision's tolerance against the FP64 oracle (split fp16: rel-L2
<= 1e-3, rel-max <= 4e-3)."""
import pytest

import oracle
from tests.helpers import complex_input, from_device, rel_errors, to_device


@pytest.mark.gpu
@pytest.mark.parametrize("shape,batch", [((65536,), 32), ((2176,), 64), ((576, 640), 2)])
def test_autotune_then_tuned_plan(dev, shape, batch):
    import fft_demo_amd as F
    from fft_demo_amd import tune

    e = tune.autotune(shape, batch, precision="split_fp16")
    print(shape, batch, e)
    assert e["candidates"] >= 1 and e["ms"] > 0
    assert e["default_ms"] < 0 or e["ms"] <= e["default_ms"]
    assert tune.lookup(shape, batch, precision="split_fp16") == e
    plan = F.Plan(shape, batch, precision="split_fp16", split_hint=tune.SPLIT_TUNED)
    if len(shape) == 1 and e["split_hint"] > 1:
        assert plan.info["passes"][0]["R"] == e["split_hint"]
    x = complex_input((batch, *shape), 21)
    y = from_device(plan.exec(to_device(x, plan.in_dtype(), dev))) / plan.out_scale
    ref = oracle.fft(x) if len(shape) == 1 else oracle.fft2d(x)
    l2, mx = rel_errors(y, ref)
    print(f"tuned {shape} b={batch}: passes={[p['R'] for p in plan.info['passes']]} rel_l2={l2:.2e} max={mx:.2e}")
    assert l2 <= 1e-3 and mx <= 4e-3
