"""batoid YAML optic descriptions -> Optic trees (batoid/parse.py).

Types are looked up in an explicit table (no ``eval``); YAML is read with
``yaml.safe_load``.  ``Clear*`` obscurations become the negation of the
matching ``Obsc*`` (parse.py:20-23)."""
import os

import numpy as np
import yaml

from .coordSys import CoordSys, RotX, RotY, RotZ
from . import surface as _surface
from . import obscuration as _obscuration
from . import medium as _medium
from . import optic as _optic

_SURFACES = {k: getattr(_surface, k) for k in (
    'Plane', 'Tilted', 'Paraboloid', 'Sphere', 'Quadric', 'Asphere', 'Zernike',
    'Bicubic')}
_OBSCS = {k: getattr(_obscuration, k) for k in (
    'ObscCircle', 'ObscAnnulus', 'ObscRay', 'ObscRectangle', 'ObscPolygon')}
_MEDIA = {k: getattr(_medium, k) for k in (
    'ConstMedium', 'TableMedium', 'SellmeierMedium', 'SumitaMedium', 'Air')}

#: directories searched by Optic.fromYaml for a bare file name
SEARCH_PATH = []


def load_yaml(filename):
    """Read ``filename``, or the first file of that name under SEARCH_PATH."""
    if os.path.exists(filename):
        with open(filename) as f:
            return yaml.safe_load(f)
    for root in SEARCH_PATH:
        for dirpath, _, files in os.walk(root):
            if os.path.basename(filename) in files:
                with open(os.path.join(dirpath, os.path.basename(filename))) as f:
                    return yaml.safe_load(f)
    raise FileNotFoundError(filename)


def parse_obscuration(config):
    config = dict(config)
    typ = config.pop('type')
    if typ in _OBSCS:
        return _OBSCS[typ](**config)
    if typ == 'ObscNegation':
        return _obscuration.ObscNegation(parse_obscuration(config['original']))
    if typ in ('ObscUnion', 'ObscIntersection'):
        items = [parse_obscuration(c) for c in config['items']]
        return getattr(_obscuration, typ)(items)
    if typ.startswith('Clear'):
        config['type'] = typ.replace("Clear", "Obsc")
        return _obscuration.ObscNegation(parse_obscuration(config))
    raise ValueError(f"Unknown obscuration type {typ}")


def parse_surface(config):
    config = dict(config)
    typ = config.pop('type')
    if typ == 'Sum':
        return _surface.Sum([parse_surface(c) for c in config['items']])
    if typ not in _SURFACES:
        raise ValueError(f"Unknown surface type {typ}")
    return _SURFACES[typ](**config)


def parse_coordSys(config, coordSys=CoordSys()):
    """Apply shifts and rotations in the order given (parse.py:37-58)."""
    for k, v in config.items():
        if k == 'x':
            coordSys = coordSys.shiftLocal([v, 0.0, 0.0])
        if k == 'y':
            coordSys = coordSys.shiftLocal([0.0, v, 0.0])
        if k == 'z':
            coordSys = coordSys.shiftLocal([0.0, 0.0, v])
        if k == 'shift':
            coordSys = coordSys.shiftLocal(v)
        if k == 'rotX':
            coordSys = coordSys.rotateLocal(RotX(v))
        if k == 'rotY':
            coordSys = coordSys.rotateLocal(RotY(v))
        if k == 'rotZ':
            coordSys = coordSys.rotateLocal(RotZ(v))
    return coordSys


def parse_medium(config):
    from numbers import Real
    if config is None:
        return None
    if isinstance(config, _medium.Medium):
        return config
    if isinstance(config, Real):
        return _medium.ConstMedium(config)
    config = dict(**config)
    typ = config.pop('type')
    if typ not in _MEDIA:
        raise ValueError(f"Unknown medium type {typ}")
    return _MEDIA[typ](**config)


_ATTRS = ('backDist', 'sphereRadius', 'pupilSize', 'pupilObscuration', 'R_outer', 'R_inner')
_INTERFACES = {
    'Mirror': _optic.Mirror, 'RefractiveInterface': _optic.RefractiveInterface,
    'Baffle': _optic.Baffle, 'Detector': _optic.Detector,
}


def parse_optic(config, coordSys=CoordSys(), inMedium=_medium.ConstMedium(1.0),
                outMedium=None):
    """Build an Optic from its configuration dict (parse.py:89-215)."""
    config = dict(config)
    obscuration = (parse_obscuration(config.pop('obscuration'))
                   if 'obscuration' in config else None)
    name = config.pop('name', "")
    if 'coordSys' in config:
        coordSys = parse_coordSys(config.pop('coordSys'), coordSys)
    inMedium = parse_medium(config.pop('inMedium', inMedium))
    outMedium = parse_medium(config.pop('outMedium', outMedium))
    if outMedium is None:
        outMedium = inMedium
    kwargs = {k: config[k] for k in _ATTRS if k in config}
    if 'stopSurface' in config:
        kwargs['stopSurface'] = parse_optic(config['stopSurface'])
    typ = config.pop('type')
    if typ in _INTERFACES:
        return _INTERFACES[typ](parse_surface(config.pop('surface')), name=name,
                                coordSys=coordSys, obscuration=obscuration,
                                inMedium=inMedium, outMedium=outMedium, **kwargs)
    if typ == 'OPDScreen':
        surface = parse_surface(config.pop('surface'))
        screen = parse_surface(config.pop('screen'))
        return _optic.OPDScreen(surface, screen, name=name, coordSys=coordSys,
                                obscuration=obscuration, inMedium=inMedium,
                                outMedium=outMedium, **kwargs)
    if typ == 'Lens':
        medium = parse_medium(config.pop('medium'))
        ic = config.pop('items')
        items = [parse_optic(ic[0], coordSys=coordSys, inMedium=inMedium, outMedium=medium),
                 parse_optic(ic[1], coordSys=coordSys, inMedium=medium, outMedium=outMedium)]
        return _optic.Lens(items, name=name, coordSys=coordSys, inMedium=inMedium,
                           outMedium=outMedium, medium=medium, **kwargs)
    if typ == 'CompoundOptic':
        items = [parse_optic(c, coordSys=coordSys, inMedium=inMedium, outMedium=outMedium)
                 for c in config.pop('items')]
        return _optic.CompoundOptic(items, inMedium=inMedium, outMedium=outMedium,
                                    name=name, coordSys=coordSys, **kwargs)
    if typ == 'Interface':
        return _optic.Interface(parse_surface(config.pop('surface')), name=name,
                                coordSys=coordSys, **kwargs)
    raise ValueError(f"Unknown optic type: {typ}")
