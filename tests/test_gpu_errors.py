"""Device-error handling, forced from the host: PH_HIP_FAIL_AFTER=n makes the
library refuse every kernel launch after the first n (before it is issued),
so no kernel ever faults.  Each case runs in a fresh child process, which
reads the variable before its first launch.  Checked: the refusal surfaces as
phase.DeviceError naming the operation and the variable, nothing computed
before it is lost, host reads (device-to-host copies, the paged mirror, a
pending fill's value) keep working, and later launches fail the same way
instead of crashing."""
import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _child(code, fail_after):
    env = dict(os.environ)
    env["PH_HIP_FAIL_AFTER"] = str(fail_after)
    env["PYTHONPATH"] = os.pathsep.join([os.path.join(ROOT, "ph-core_amd"), os.path.join(ROOT, "tests"),
                                         env.get("PYTHONPATH", "")])
    print(f"child with PH_HIP_FAIL_AFTER={fail_after}", flush=True)
    r = subprocess.run([sys.executable, "-c", textwrap.dedent(code)], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_every_launch_refused():
    out = _child("""
        import numpy as np
        import phase
        phase.init(0)
        h = np.arange(1000, dtype=np.int32)
        x = phase.NArray.from_numpy(h)          # host-to-device copy: no launch
        assert np.array_equal(x.to_numpy(), h)  # device-to-host copy: no launch
        for op in (lambda: (x + 1).materialize(), lambda: x.sum(), lambda: x.view("..2..").to_narr(),
                   lambda: x.process(lambda v: v * 2).to_narr().materialize(),
                   lambda: (x.wadd(x) * 1.5).materialize()):
            try:
                op()
                raise SystemExit("a refused launch did not raise")
            except phase.DeviceError as e:
                assert "PH_HIP_FAIL_AFTER" in str(e), str(e)
        f = phase.NArray.fill([64, 64], 7)      # pending: no launch yet
        assert f.get([3, 5]) == 7               # answered from the pending value
        try:
            f.to_numpy()                        # needs the fill kernel
            raise SystemExit("a refused fill did not raise")
        except phase.DeviceError:
            pass
        assert f.get([63, 0]) == 7
        assert np.array_equal(x.to_numpy(), h)  # earlier data intact
        print("refused cleanly")
    """, 0)
    assert "refused cleanly" in out


def test_refusal_after_some_launches_keeps_earlier_results():
    out = _child("""
        import numpy as np
        import phase
        phase.init(0)
        h = np.arange(4096, dtype=np.float64).reshape(64, 64)
        x = phase.NArray.from_numpy(h)
        done = []
        try:
            for k in range(100000):
                y = (x * float(k) + 1.0).materialize()
                done.append((k, y))
        except phase.DeviceError as e:
            assert "PH_HIP_FAIL_AFTER=40" in str(e), str(e)
        assert 5 <= len(done) <= 40, len(done)
        for k, y in done:                        # every finished result is whole
            assert np.array_equal(y.to_numpy(), h * float(k) + 1.0), k
        try:
            x.sum()
            raise SystemExit("no refusal after the limit")
        except phase.DeviceError:
            pass
        print("kept", len(done))
    """, 40)
    assert "kept" in out
