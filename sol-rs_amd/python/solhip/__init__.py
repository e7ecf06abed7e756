"""solhip: Python binding of the MI355X ray-tracing backend (libsolhip.so).

`api` binds the C ABI of include/solhip.h; `scenes` builds seeded synthetic
scenes with the reference's record layouts; `records` holds those layouts.
"""
from . import records, scenes  # noqa: F401
from .api import (BUILD_PREFER_FAST_TRACE, PROG_AO, PROG_AO_GLSL, PROG_DEBUG, PROG_PATHTRACE, PROG_PATHTRACE_GLSL, TRACE_COUNT,  # noqa: F401
                  JOB_DONE, JOB_QUEUED, JOB_RUNNING, TRACE_TERMINATE_ON_FIRST_HIT, Context, GpuScene, Images,
                  Mlp, Server, SolhipError, Spec, lib, spec_default)
