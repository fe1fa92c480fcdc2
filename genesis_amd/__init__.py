"""MI355X-native batched rigid-body step with the gs.Scene / RigidEntity surface.

    import genesis_amd as gs
    gs.init(backend=gs.gpu)
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01))
    scene.add_entity(gs.morphs.Plane())
    arm = scene.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml"))
    scene.build(n_envs=64)
    arm.control_dofs_position(targets)
    scene.step()
"""
from .model import (CTRL_FORCE, CTRL_POSITION, CTRL_VELOCITY, GEOM_BOX, GEOM_CAPSULE, GEOM_PLANE, GEOM_SPHERE,
                    JOINT_FIXED, JOINT_FREE, JOINT_PRISMATIC, JOINT_REVOLUTE, JOINT_SPHERICAL)
from .scene import RigidEntity, Scene, materials, morphs, options  # noqa: F401

__version__ = "0.1.0"

gpu = "gpu"
cpu = "cpu"
_initialized = False
backend = None
precision = "32"


class GEOM_TYPE:  # noqa: N801 - values of genesis/constants.py:19-29
    PLANE, SPHERE, ELLIPSOID, CYLINDER, CAPSULE, BOX, MESH, TERRAIN = range(8)


class JOINT_TYPE:  # noqa: N801
    FIXED, REVOLUTE, PRISMATIC, SPHERICAL, FREE = range(5)


class CTRL_MODE:  # noqa: N801
    FORCE, VELOCITY, POSITION = range(3)


def init(backend=gpu, precision="32", seed=None, logging_level=None, **_ignored):
    """Select the device. Only the GPU backend in FP32 is provided: the step is HIP code for gfx950."""
    import torch
    g = globals()
    if backend not in (gpu, "cuda", "amdgpu"):
        raise NotImplementedError("the rigid step runs on the GPU (HIP, gfx950) only")
    if str(precision) != "32":
        raise NotImplementedError("the HIP step computes in FP32")
    if seed is not None:
        torch.manual_seed(seed)
    g["backend"], g["precision"], g["_initialized"] = gpu, "32", True


def destroy():
    globals()["_initialized"] = False
