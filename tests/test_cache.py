"""The build-artefact cache (kinetix_amd/cache.py) in the reference's on-disk layout
(benchmark/src/kinetix.cpp:318-334, 677-726): .cache/KINETIX/<yaml stem>/ with the
reference-named routine files at the top and .hash holding the decimal 31-bit hash of the
generator command line; regeneration only when that hash (or an artefact) is stale. CPU only.

The hash is pinned against the reference's expression compiled with its own C types
(kinetix.cpp:111-121: unsigned int state, a char -- signed on x86 -- per step), which is where a
restatement in Python can go wrong (bytes >= 0x80)."""
import ctypes as C
import os
import random
import shutil
import subprocess

import pytest

from conftest import ROOT
from kinetix_amd import api, cache

YAML = os.path.join(ROOT, 'kinetix_amd', 'mechs', 'synth40cov', 'synth40cov.yaml')

HASH_C = r'''
#include <stddef.h>
/* kinetix.cpp:111-121, with the reference's types */
unsigned long ref_hash(const char* str, size_t n)
{
  unsigned int hash = 1315423911;
  for (size_t i = 0; i < n; i++) hash ^= ((hash << 5) + str[i] + (hash >> 2));
  return (hash & 0x7FFFFFFF);
}
'''


def test_hash_matches_the_reference_expression(tmp_path):
    src = tmp_path / 'h.c'
    src.write_text(HASH_C)
    so = tmp_path / 'libh.so'
    subprocess.run(['gcc', '-O1', '-shared', '-fPIC', str(src), '-o', str(so)], check=True)
    L = C.CDLL(str(so))
    L.ref_hash.restype = C.c_ulong
    L.ref_hash.argtypes = [C.c_char_p, C.c_size_t]
    rng = random.Random(7)
    cases = ['', 'a', '/opt/kinetix/kinetix/__main__.py --header-only --mechanism gri30.yaml',
             'Grün éè ☃ mech.yaml']
    cases += [''.join(chr(rng.randrange(1, 0x3000)) for _ in range(rng.randrange(1, 300)))
              for _ in range(200)]
    for s in cases:
        b = s.encode()
        assert cache.js_hash(s) == L.ref_hash(b, len(b)), s
    assert cache.js_hash('') == 1315423911


def _mtime(d, f):
    return os.stat(os.path.join(d, f)).st_mtime_ns


def test_layout_reuse_and_invalidation(tmp_path):
    y = tmp_path / 'mymech.yaml'
    shutil.copyfile(YAML, y)
    root = str(tmp_path / 'cache')
    d = cache.ensure(str(y), root=root, build=False)
    assert d == os.path.join(root, 'mymech')
    for f in cache.artefacts(build=False):
        assert os.path.isfile(os.path.join(d, f)), f
    # the reference-named files at the top are the generated routine files
    for f in cache.ROUTINE_FILES + cache.TRANSPORT_FILES:
        assert open(os.path.join(d, f)).read() == open(os.path.join(d, 'routines', f)).read()
    h = cache.js_hash(cache.command_line(str(y), d))
    assert open(os.path.join(d, '.hash')).read() == str(h)          # decimal, as the reference
    t0 = _mtime(d, 'kernels.hip')
    # same command line: nothing regenerated
    assert cache.ensure(str(y), root=root, build=False) == d
    assert _mtime(d, 'kernels.hip') == t0
    # an edited mechanism file under the same path regenerates
    with open(y, 'a') as fh:
        fh.write('\n# edited\n')
    cache.ensure(str(y), root=root, build=False)
    t1 = _mtime(d, 'kernels.hip')
    assert t1 != t0 and cache.stored_hash(d) != h
    # other generator flags: another hash, regenerated
    cache.ensure(str(y), flags=('--rolled',), root=root, build=False)
    assert _mtime(d, 'kernels.hip') != t1
    assert 'kx_bk1_fp64' in open(os.path.join(d, 'kernels.hip')).read()
    h_rolled = cache.stored_hash(d)
    assert h_rolled == cache.js_hash(cache.command_line(str(y), d, ('--rolled',)))
    # a corrupt .hash or a missing artefact: regenerated
    open(os.path.join(d, '.hash'), 'w').write('garbage')
    cache.ensure(str(y), flags=('--rolled',), root=root, build=False)
    assert cache.stored_hash(d) == h_rolled
    os.remove(os.path.join(d, 'mech.h'))
    cache.ensure(str(y), flags=('--rolled',), root=root, build=False)
    assert os.path.isfile(os.path.join(d, 'mech.h'))
    with pytest.raises(ValueError):
        cache.ensure(str(y), flags=('--unroll',), root=root, build=False)
    with pytest.raises(FileNotFoundError):
        cache.ensure(str(tmp_path / 'absent.yaml'), root=root)


def test_yaml_path_builds_the_code_object_once(tmp_path, monkeypatch):
    """Kinetix('<file>.yaml') resolves through the cache: generated and compiled for gfx950 on
    the first call (the reference's init), reused on the next."""
    y = tmp_path / 'cov.yaml'
    shutil.copyfile(YAML, y)
    monkeypatch.setenv('KINETIX_CACHE', str(tmp_path / 'kc'))
    d = api.mech_dir(str(y))
    assert d == os.path.join(str(tmp_path / 'kc'), 'cov')
    assert os.path.getsize(os.path.join(d, 'kernels.hsaco')) > 0
    t = _mtime(d, 'kernels.hsaco')
    assert api.mech_dir(str(y)) == d and _mtime(d, 'kernels.hsaco') == t
