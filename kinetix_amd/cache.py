"""Build-artefact cache in the reference's on-disk layout.

The reference's init (benchmark/src/kinetix.cpp:318-334, 677-726) generates a mechanism into
`.cache/KINETIX/<yaml stem>/` and keeps a file `.hash` there: the decimal 31-bit hash
(kinetix.cpp:111-121) of the generator command line. The generator runs again only when the
stored number differs, and a fresh generation forces a kernel rebuild (the `.occa.bck`
rotation, kinetix.cpp:719-726).

`ensure(yaml)` keeps the same directory and `.hash` file for this generator's output:
  * mech.h, rates.cpp, enthalpy_RT.cpp, heat_capacity_R.cpp (and, with transport,
    conductivity.cpp, viscosity.cpp, diffusivity.cpp) at the top, the reference's file names
    (also under routines/, where kernels.hip includes them from);
  * tables.npz, mech.json, mech.txt, counts.json, kernels.hip and the gfx950 code object
    kernels.hsaco, built ahead of time (this replaces OCCA's kernel cache): the directory is
    what `Kinetix(dir)` / `kx_init` load.

Differences from the reference, each on purpose:
  * the hashed command line also names a digest of the generator's own sources and the
    SHA-256 of the YAML file, so a newer generator or an edited mechanism file regenerates
    (the reference hashes its command line alone, so an edit under the same path is missed);
  * `.hash` is written last, after a complete generation and build, and removed before a
    regeneration starts: an interrupted build leaves the entry stale, never half-valid;
  * an exclusive lock on `.lock` serialises processes that share a cache (the reference
    generates on rank 0 behind an MPI barrier, kinetix.cpp:672-730).
"""
import contextlib
import fcntl
import glob
import hashlib
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_ROOT = os.path.join('.cache', 'KINETIX')      # relative to the working directory
ROUTINE_FILES = ('mech.h', 'rates.cpp', 'enthalpy_RT.cpp', 'heat_capacity_R.cpp')
TRANSPORT_FILES = ('conductivity.cpp', 'viscosity.cpp', 'diffusivity.cpp')
FLAGS = ('--fit-rcpdiffcoeffs', '--rolled', '--no-transport')


def js_hash(text):
    """The reference's string hash (kinetix.cpp:111-121): h = 1315423911, then per char
    h ^= (h << 5) + c + (h >> 2) in 32-bit unsigned arithmetic, with c a *signed* char (x86),
    result h & 0x7FFFFFFF."""
    h = 1315423911
    for b in text.encode():
        c = b - 256 if b >= 128 else b
        h ^= ((h << 5) + c + (h >> 2)) & 0xFFFFFFFF
    return h & 0x7FFFFFFF


def generator_digest():
    """SHA-256 over the generator's sources (the code that decides what is generated and how
    it compiles): any change to them invalidates every cache entry."""
    files = sorted(glob.glob(os.path.join(PKG, 'gen', '*.py')))
    files += [os.path.join(PKG, 'build.py'), os.path.join(PKG, 'csrc', 'kx_mlp_wave.h'),
              os.path.join(PKG, 'csrc', 'kx_rcp.h')]
    h = hashlib.sha256()
    for f in files:
        h.update(os.path.basename(f).encode() + b'\0')
        with open(f, 'rb') as fh:
            h.update(fh.read())
    return h.hexdigest()


def cache_dir(yaml_path, root=None):
    """<root>/<yaml stem>, absolute; root: argument, else $KINETIX_CACHE, else .cache/KINETIX
    under the working directory (the reference's location)."""
    root = root or os.environ.get('KINETIX_CACHE') or DEFAULT_ROOT
    stem = os.path.splitext(os.path.basename(yaml_path))[0]
    return os.path.join(os.path.abspath(root), stem)


def command_line(yaml_path, out_dir, flags=()):
    """The hashed string: the generator invocation, as the reference hashes its own
    (kinetix.cpp:669-677), plus the generator and mechanism digests."""
    with open(yaml_path, 'rb') as fh:
        mech_sha = hashlib.sha256(fh.read()).hexdigest()
    cmd = f'{PKG}/gen --mechanism {os.path.abspath(yaml_path)} --output {out_dir}'
    cmd += ''.join(f' {f}' for f in flags)
    return cmd + f' --generator {generator_digest()} --mechanism-sha256 {mech_sha}'


def stored_hash(d):
    """The number in d/.hash, or None (absent or unreadable: the reference reads 0 then)."""
    try:
        with open(os.path.join(d, '.hash')) as fh:
            return int(fh.read().split()[0])
    except (OSError, ValueError, IndexError):
        return None


def artefacts(flags=(), build=True):
    """The files a complete entry holds for these flags."""
    need = ('tables.npz', 'mech.json', 'mech.txt', 'kernels.hip') + ROUTINE_FILES
    if '--no-transport' not in flags:
        need += TRANSPORT_FILES
    return need + (('kernels.hsaco',) if build else ())


def _complete(d, flags, build):
    return all(os.path.isfile(os.path.join(d, f)) for f in artefacts(flags, build))


@contextlib.contextmanager
def _locked(d):
    with open(os.path.join(d, '.lock'), 'a') as fh:
        fcntl.flock(fh, fcntl.LOCK_EX)
        try:
            yield
        finally:
            fcntl.flock(fh, fcntl.LOCK_UN)


def ensure(yaml_path, flags=(), root=None, build=True, log=None):
    """Directory of yaml_path's generated (and, with build, compiled) artefacts, regenerating
    only when the stored hash differs from this command line's or an artefact is missing.
    Returns the directory."""
    flags = tuple(flags)
    bad = [f for f in flags if f not in FLAGS]
    if bad:
        raise ValueError(f'unknown generator flags {bad}; known: {FLAGS}')
    if not os.path.isfile(yaml_path):
        raise FileNotFoundError(yaml_path)
    d = cache_dir(yaml_path, root)
    os.makedirs(d, exist_ok=True)
    h = js_hash(command_line(yaml_path, d, flags))
    with _locked(d):
        if stored_hash(d) == h and _complete(d, flags, build):
            return d
        for f in ('.hash', 'kernels.hsaco') + ROUTINE_FILES + TRANSPORT_FILES:
            with contextlib.suppress(FileNotFoundError):
                os.remove(os.path.join(d, f))
        if log:
            log(f'kinetix cache: generating {os.path.basename(d)} in {d}')
        name = os.path.basename(d)
        env = dict(os.environ)
        env['PYTHONPATH'] = os.pathsep.join([os.path.dirname(PKG)] +
                                            ([env['PYTHONPATH']] if env.get('PYTHONPATH') else []))
        subprocess.run([sys.executable, '-m', 'kinetix_amd.gen', '--mechanism',
                        os.path.abspath(yaml_path), '--output', d, '--name', name, *flags],
                       check=True, env=env, stdout=subprocess.DEVNULL)
        for f in ROUTINE_FILES + TRANSPORT_FILES:        # the reference's layout: at the top
            if os.path.isfile(os.path.join(d, 'routines', f)):
                shutil.copyfile(os.path.join(d, 'routines', f), os.path.join(d, f))
        if build:
            from . import build as _build
            _build.build_mechanism(d)
        tmp = os.path.join(d, '.hash.tmp')
        with open(tmp, 'w') as fh:
            fh.write(f'{h}')
        os.replace(tmp, os.path.join(d, '.hash'))
    return d
