"""The one-call neighbour search (cusoap_nl_build) and its two fallbacks.

SOAP keeps the buffer sizes of its last search per system shape and builds the next
list with one call. The lists must be bit-identical to the three-call search in every
case:
- the sizes fit;
- the planted workspace is too small, so the call returns CUSOAP_ERR_WORKSPACE and the
  count runs with the plan's size;
- the planted edge capacity is too small, so the fill is skipped and cusoap_nl_fill
  completes it.
"""

import sys

import numpy as np
import pytest
import torch

from cusoap_amd import SOAP, systems

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lists(nl):
    return nl.offsets.cpu().numpy(), nl.nbr[:nl.n_edges].cpu().numpy(), nl.max_count


@pytest.mark.parametrize("periodic", [True, False])
def test_nl_build_paths_give_the_same_lists(periodic):
    if periodic:
        s = systems.water_box(6, seed=31)
    else:
        s = systems.water_cluster_300(seed=31)
    g = SOAP(device=DEV, dtype="float64", r_cut=5.0, n_max=2, l_max=1, sigma=0.5, species=[1, 8], periodic=periodic)
    x = g._context(s, None)
    key = (x.n_centers, x.n_atoms, x.n_img, g._cutoff, False)
    g._nl_sizes.clear()
    ref = _lists(g._neighbors(x, drop_self=False))            # three calls; stores the sizes
    assert key in g._nl_sizes
    ws_bytes, room = g._nl_sizes[key]
    cases = {"sizes fit": (ws_bytes, room), "workspace too small": (256, room), "edges over capacity": (ws_bytes, 1)}
    for name, sizes in cases.items():
        g._nl_sizes[key] = sizes
        got = _lists(g._neighbors(x, drop_self=False))
        same = all(np.array_equal(a, b) for a, b in zip(got[:2], ref[:2])) and got[2] == ref[2]
        say(f"nl_build periodic={periodic} {name}: {ref[1].size} edges, identical {same}")
        assert same
        assert g._nl_sizes[key][0] >= sizes[0] and g._nl_sizes[key][1] >= ref[1].size
