"""Writes the golden vectors in this folder: small workloads traced by the
numpy oracle, stored as inputs, pre-drawn uniforms and every output column
(``<case>.npz``) plus the plain-data instrument description they were traced
with and how it was built (``<case>.json``).

    python tests/golden/make_golden.py [case ...]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from helpers import encode_description, make_inputs  # noqa: E402
from oracle import trace as OT  # noqa: E402
from oracle.describe import describe  # noqa: E402

CASES = {
    # name: (builder, builder args, photons, seed, energies)
    'flat_grid_1000': ('flat_grid_instrument', {}, 1000, 31, 'uniform'),
    'cat_rowland_200_2000': ('cat_rowland_instrument', {'n_facets': 200}, 2000, 32, 'lines'),
    'cat_stack_200_2000': ('cat_stack_instrument', {'n_facets': 200}, 2000, 33, 'lines'),
}


def build(builder, kwargs):
    from marxs_amd import configs
    inst = getattr(configs, builder)(**kwargs)
    return inst[0] if isinstance(inst, tuple) else inst


def main(names=None):
    for name, (builder, kwargs, n, seed, energies) in CASES.items():
        if names and name not in names:
            continue
        inst = build(builder, kwargs)
        els = describe(inst)
        n_slots = OT.uniform_slots(els)[1]
        ph, u = make_inputs(n, seed, n_slots, energies)
        out = OT.trace(els, ph, u)
        arrays = {'uniforms': u}
        arrays.update({'in_' + k: v for k, v in ph.items()})
        arrays.update({'out_' + k: np.asarray(v) for k, v in out.items()})
        np.savez_compressed(os.path.join(HERE, name + '.npz'), **arrays)
        meta = {'builder': builder, 'kwargs': kwargs, 'photons': n, 'seed': seed,
                'energies': energies, 'n_slots': int(n_slots),
                'description': encode_description(els)}
        with open(os.path.join(HERE, name + '.json'), 'w') as f:
            json.dump(meta, f)
        print(name, os.path.getsize(os.path.join(HERE, name + '.npz')),
              os.path.getsize(os.path.join(HERE, name + '.json')))


if __name__ == '__main__':
    main(sys.argv[1:])
