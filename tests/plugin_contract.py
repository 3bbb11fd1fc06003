"""Resolution of a fill backend the way pcbqs finds plugins: by name in the entry-point group
`pcbqs.static_backends`, exactly one provider, loaded lazily, the factory called as
factory(geometry, workers, chunk, device) (the contract of src/pcbqs/static_backends.py:253-271
and docs/static-backends.md:157-193).

The package is not installed here (it runs from its source tree), so `entry_point_path` writes
a throw-away distribution (a .dist-info folder holding entry_points.txt) made from the
entry points that pyproject.toml declares; importlib.metadata then finds it on sys.path as it
would find an installed distribution.
"""
import os
import sys
from importlib import metadata

try:
    import tomllib
except ImportError:                                           # Python 3.10
    import tomli as tomllib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GROUP = 'pcbqs.static_backends'


def declared_entry_points():
    with open(os.path.join(ROOT, 'pyproject.toml'), 'rb') as fh:
        project = tomllib.load(fh)['project']
    return project['name'], project['version'], project.get('entry-points', {})


def entry_point_path(tmp_path):
    """Write <name>-<version>.dist-info with pyproject.toml's entry points under tmp_path and put
    tmp_path on sys.path (the caller removes it)."""
    name, version, groups = declared_entry_points()
    info = tmp_path / f"{name.replace('-', '_')}-{version}.dist-info"
    info.mkdir()
    (info / 'METADATA').write_text(f'Metadata-Version: 2.1\nName: {name}\nVersion: {version}\n')
    lines = []
    for group, entries in groups.items():
        lines.append(f'[{group}]')
        lines += [f'{key} = {value}' for key, value in entries.items()]
    (info / 'entry_points.txt').write_text('\n'.join(lines) + '\n')
    sys.path.insert(0, str(tmp_path))
    return str(tmp_path)


def resolve(name):
    """The one factory registered under `name` in the group; ValueError when none or several."""
    found = [ep for ep in metadata.entry_points(group=GROUP) if ep.name == name]
    if len(found) != 1:
        raise ValueError(f'fill backend {name!r}: {len(found)} providers in {GROUP}')
    return found[0].load()
