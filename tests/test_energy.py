"""CPU tests of the CU-policy energy harness (the GPU counterpart of the
reference's per-task-type DVFS, palial_power/server.c:69-84): the policy
parser the library uses (palial_hip_cu_policy), the hwmon sensor lookup on a
fake sysfs tree, and the energy integral on samples with a known answer."""
import ctypes as C

import pytest

from palial_amd import energy


def policy(text, ncu=256):
    import palial_amd
    L = palial_amd.lib()
    L.palial_hip_cu_policy.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_uint), C.c_int]
    words = (ncu + 31) // 32
    buf = (C.c_uint * (3 * words))()
    rc = L.palial_hip_cu_policy(text.encode(), ncu, buf, words)
    if rc < 0:
        return rc, None
    sets = []
    for c in range(3):
        cus = {w * 32 + bit for w in range(words) for bit in range(32) if buf[c * words + w] >> bit & 1}
        sets.append(cus)
    return rc, sets


def test_policy_parse():
    rc, (bulk, aux, chain) = policy("bulk=8-255,aux=8-71,chain=0-7")
    assert rc == 1 and bulk == set(range(8, 256)) and aux == set(range(8, 72)) and chain == set(range(8))
    rc, (bulk, aux, chain) = policy("bulk=8-191")
    assert rc == 0 and bulk == set(range(8, 192)) and aux == bulk and chain == set()
    rc, (bulk, aux, chain) = policy("aux=3+5-6+100", ncu=104)
    assert bulk == set(range(104)) and aux == {3, 5, 6, 100}
    for name in energy.POLICIES.values():
        if "PALIAL_CU_POLICY" in name:
            assert policy(name["PALIAL_CU_POLICY"])[0] >= 0


@pytest.mark.parametrize("bad", ["", "bulk", "bulk=", "bulk=5-3", "bulk=0-256", "gemm=0-7", "bulk=1,bulk=2",
                                 "bulk=1;aux=2", "bulk=-1", "chain=0-7,", "bulk=1-"])
def test_policy_rejects(bad):
    assert policy(bad)[0] < 0


def test_sensor_lookup(tmp_path):
    assert energy.find_power_sensor(str(tmp_path)) is None
    mon = tmp_path / "card1" / "device" / "hwmon" / "hwmon4"
    mon.mkdir(parents=True)
    (mon / "name").write_text("nvme\n")
    (mon / "power1_average").write_text("1000\n")
    assert energy.find_power_sensor(str(tmp_path)) is None          # not an amdgpu sensor
    (mon / "name").write_text("amdgpu\n")
    (mon / "power1_input").write_text("523000000\n")
    assert energy.find_power_sensor(str(tmp_path)).endswith("power1_average")
    (mon / "power1_average").write_text("n/a\n")
    assert energy.find_power_sensor(str(tmp_path)).endswith("power1_input")


def test_integrate():
    # P(t) = 100 + 10 t, sampled every 0.1 s on [0, 2]: exact for a linear power
    s = [(0.1 * i, 100 + i) for i in range(21)]
    e, w = energy.integrate(s, 0.25, 1.75)
    exact = 100 * 1.5 + 5 * (1.75 ** 2 - 0.25 ** 2)
    assert abs(e - exact) < 1e-9 and abs(w - exact / 1.5) < 1e-9
    assert energy.integrate(s, -1.0, 1.0) == (None, None)            # window not covered
    assert energy.integrate(s[:1], 0.0, 0.0) == (None, None)
