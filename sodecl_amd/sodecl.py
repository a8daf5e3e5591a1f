"""The reference's Python wrapper (interfaces/python/sodecl.py:35-84 of
avramidis/sodecl), same signature and result contract, over the MI355X
sodecl_interface module.

    t, results = sodecl(openclplatform, opencldevice, openclkernel,
                        initx, params, sodesolver, orbits, nequat,
                        nnoiseprocesses, dt, tspan, ksteps, localgroupsize)

results has shape (orbits*nequat, nsnap) with row eq*orbits + orbit;
t = arange(0, tspan + dt*ksteps, dt*ksteps) (its length can differ from
nsnap exactly as in the reference).  openclplatform must be 0 and
opencldevice is the HIP device ordinal.  openclkernel is the system source
path; a bare name of a bundled system (e.g. 'oculomotor.cl') that does not
exist in the working directory resolves to sodecl_amd/systems/.
"""
import os

import numpy

from . import sodecl_interface

_SYSTEMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "systems")


def _resolve(path):
    if os.path.exists(path):
        return path
    alt = os.path.join(_SYSTEMS, os.path.basename(path))
    return alt if os.path.exists(alt) else path


def sodecl(openclplatform, opencldevice, openclkernel,
           initx, params, sodesolver,
           orbits, nequat, nnoiseprocesses,
           dt, tspan, ksteps, localgroupsize):
    t0 = numpy.zeros(orbits)                       # sodecl.py:40-44
    nparams = numpy.asarray(params)[0].size        # sodecl.py:42
    initx = numpy.asarray(initx, dtype=numpy.float64).flatten()
    params = numpy.asarray(params, dtype=numpy.float64).flatten()

    # solver forcing (sodecl.py:48-61)
    if nnoiseprocesses == 0:
        if 0 < sodesolver < 5:
            solver2user = sodesolver
        else:
            print("Invalid SODE solver selected! Using Euler!")
            solver2user = 1
    else:
        solver2user = 0

    results = sodecl_interface.sodeclcall(t0, initx, params, openclplatform, opencldevice,
                                          _resolve(openclkernel), solver2user, orbits, nequat,
                                          nparams, nnoiseprocesses, dt, tspan, ksteps,
                                          localgroupsize)
    results = numpy.asarray(results)
    results = results.reshape(orbits * nequat, int(results.shape[0] / (orbits * nequat)), order='F')
    t = numpy.arange(0, tspan + dt * ksteps, dt * ksteps)
    return t, results
