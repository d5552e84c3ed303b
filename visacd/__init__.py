"""Drop-in module with the reference's Python API (``import visacd``).

Same names and call signatures as the reference's pybind module
(lib/bindings.cpp:14-70): ``process(mesh, concavity, num_parts)``,
``config``, ``Config``, ``set_seed``, ``Mesh``, ``MeshList``, ``VecArray3d``,
``VecArray3i``, ``make_vecarray3i`` and ``ProcessResult``.  ``engine`` holds
the options of this engine that the reference does not have; its
``proposal_weights`` starts at the trained proposal model shipped with the
package (used when ``engine.proposal = "learned"``).
"""
import os as _os

from visacd_amd import share_hip_runtime

share_hip_runtime()

from visacd_amd._visacd import (  # noqa: E402,F401
    Config,
    EngineOptions,
    Mesh,
    MeshList,
    ProcessResult,
    VecArray3d,
    VecArray3i,
    config,
    engine,
    make_vecarray3i,
    process,
    set_seed,
)

_weights = _os.path.join(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))),
                         "visacd_amd", "data", "proposal_mlp.f32")
if not engine.proposal_weights and _os.path.exists(_weights):
    engine.proposal_weights = _weights
