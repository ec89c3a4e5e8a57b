"""The FP32 slab pre-test never contradicts the reference's FP64 slab test (bvh.ts:340-356): a
CPU restatement of its arithmetic (tests/native/slab_pretest_check.cpp) against the FP64 test on
random and deliberately grazing boxes. The GPU kernels are checked against the oracle by the
any-hit and render parity tests."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_slab_pretest_agrees_with_fp64(tmp_path):
    exe = tmp_path / 'slab_check'
    src = os.path.join(HERE, 'native', 'slab_pretest_check.cpp')
    r = subprocess.run(['g++', '-O2', '-std=c++17', '-ffp-contract=off', '-o', str(exe), src],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip(f'g++ unavailable: {r.stderr[:200]}')
    r = subprocess.run([str(exe), '3000000'], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert 'bad 0 ' in r.stdout and r.stdout.count('bad 0') == 2, r.stdout
