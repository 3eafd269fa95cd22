"""simuse_net_raster_stats_noalloc against simuse_net_raster_stats (needs an MI355X).

The variant that allocates nothing must give the existing entry point's
results bit for bit: two calls in a row, the second with a smaller input on
another network, plus the size limit and the empty call.
"""
import numpy as np
import pytest

from simuse import SimuseError
from simuse._lib import E_INVALID, E_STATE, RASTER_STATS_NOALLOC_MAX
from simuse.workloads import build_device, izhikevich_network
from nets import random_net

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _recorded(spec, steps):
    dev = build_device(spec, graph_steps=16)
    dev.record(steps)
    dev.run(steps)
    return dev


def test_raster_stats_noalloc_equals_allocating_entry_point():
    print("noalloc: izhikevich 80k (two populations), then LIF 5k", flush=True)
    big = _recorded(izhikevich_network(n=80_000, k=20, seed=3, steps=600), 600)
    small = _recorded(random_net(5000, 50, seed=5, model="lif", max_delay=5, noise=3.0), 800)
    # first call: the largest range allowed, across the population boundary at 64000
    first, count = 10_000, RASTER_STATS_NOALLOC_MAX
    r_ref, c_ref = big.raster_stats(first, count, 1e-3)
    r_na, c_na = big.raster_stats_noalloc(first, count, 1e-3)
    assert np.array_equal(_bits(r_na), _bits(r_ref)) and np.array_equal(_bits(c_na), _bits(c_ref))
    assert np.count_nonzero(r_ref) > 1000 and np.count_nonzero(c_ref) > 100
    # second call, smaller, on the other network: nothing of the first call leaks
    first, count = 1234, 2000
    r_ref, c_ref = small.raster_stats(first, count, 1e-4)
    r_na, c_na = small.raster_stats_noalloc(first, count, 1e-4)
    assert np.array_equal(_bits(r_na), _bits(r_ref)) and np.array_equal(_bits(c_na), _bits(c_ref))
    assert np.count_nonzero(r_ref) > 100
    # limits
    with pytest.raises(SimuseError) as e:
        big.raster_stats_noalloc(0, RASTER_STATS_NOALLOC_MAX + 1, 1e-3)
    assert e.value.code == E_INVALID
    with pytest.raises(SimuseError) as e:
        small.raster_stats_noalloc(4000, 1001, 1e-4)
    assert e.value.code == E_INVALID
    r0, c0 = small.raster_stats_noalloc(0, 0, 1e-4)
    assert r0.size == 0 and c0.size == 0
    big.close()
    small.close()
    off = build_device(random_net(300, 10, seed=1, model="lif"), graph_steps=1)
    with pytest.raises(SimuseError) as e:
        off.raster_stats_noalloc(0, 10, 1e-4)
    assert e.value.code == E_STATE
    off.close()
