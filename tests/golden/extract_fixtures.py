"""Extract the reference test fixtures into plain numpy archives.

Reads ``test/fixtures.py`` of the reference as *text*, parses it with
``ast`` and converts the three dictionary literals (``dtemp``) that the
fixtures hand to ``xr.Dataset.from_dict`` into arrays.  Nothing from the
reference is imported or executed: the only non-literal node the walker
accepts is the attribute ``np.nan``.

Outputs (data only, committed next to this script):

* ``era5_pressure.npz``  - 17 ERA5 columns x 37 pressure levels
  (reference ``test/fixtures.py:29-162``): 3-D ``t, td, u, v, q, p3d`` with
  dims (longitude, level), 1-D ``level`` (hPa), surface ``p, t, td, u, v,
  q, start3d`` and the CAPE/CIN/SRH goldens.
* ``soundings.npz`` - 13 soundings x 108 levels (``test/fixtures.py:1165-3441``),
  stored *after* the ``ds.transpose()`` at ``fixtures.py:3440``, i.e. dims
  (n2, level) with level 0 the surface, plus the CAPE/CIN/MU/SRH goldens.

Usage: ``python tests/golden/extract_fixtures.py [/root/reference]``.
"""
import ast
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _eval(node):
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_eval(node.operand)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.UAdd):
        return +_eval(node.operand)
    if isinstance(node, (ast.List, ast.Tuple)):
        return [_eval(e) for e in node.elts]
    if isinstance(node, ast.Dict):
        return {_eval(k): _eval(v) for k, v in zip(node.keys, node.values)}
    if (isinstance(node, ast.Attribute) and node.attr == 'nan'
            and isinstance(node.value, ast.Name) and node.value.id == 'np'):
        return float('nan')
    raise ValueError(f'unsupported node {ast.dump(node)[:80]}')


def load_literals(path):
    """Return {function name: [dtemp dicts in order]} from the fixture file."""
    tree = ast.parse(open(path).read())
    out = {}
    for fn in tree.body:
        if not isinstance(fn, ast.FunctionDef):
            continue
        for stmt in fn.body:
            if (isinstance(stmt, ast.Assign) and len(stmt.targets) == 1
                    and isinstance(stmt.targets[0], ast.Name)
                    and stmt.targets[0].id == 'dtemp'):
                out.setdefault(fn.name, []).append(_eval(stmt.value))
    return out


def _vars(d):
    res = {}
    for name, spec in d['data_vars'].items():
        res[name] = (tuple(spec['dims']), np.asarray(spec['data'], dtype=np.float64))
    for name, spec in d['coords'].items():
        res['coord_' + name] = (tuple(spec['dims']), np.asarray(spec['data'], dtype=np.float64))
    return res


def main(ref_root):
    lits = load_literals(os.path.join(ref_root, 'test', 'fixtures.py'))

    # ERA5: first dict is the surface dataset, second the 3-D dataset
    # (fixtures.py:150 and :1161).
    surf, d3 = (_vars(d) for d in lits['dataset_ERA5pressurelevel'])
    era = {}
    for k, (dims, a) in surf.items():
        if k.startswith('coord_'):
            continue
        era['sfc_' + k] = a
    for k, (dims, a) in d3.items():
        if k == 'coord_level':
            era['level'] = a
            continue
        if k.startswith('coord_'):
            continue
        assert dims == ('longitude', 'level'), (k, dims)
        era[k] = a
    np.savez(os.path.join(HERE, 'era5_pressure.npz'), **era)

    (snd,) = lits['dataset_soundings']
    sv = _vars(snd)
    out = {}
    for k, (dims, a) in sv.items():
        if k.startswith('coord_'):
            continue
        # ds.transpose() reverses every variable's dims (fixtures.py:3440)
        out[k] = a.T.copy() if a.ndim == 2 else a
    np.savez(os.path.join(HERE, 'soundings.npz'), **out)
    for name, d in (('era5_pressure', era), ('soundings', out)):
        print(name, {k: v.shape for k, v in d.items()})


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
