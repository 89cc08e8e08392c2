"""pyticles_amd -- MI355X (gfx950) smooth-particle hot path.

Hand-written HIP kernels for the periodic cell-list neighbour build, Lucy
summation density with the van der Waals EOS, and the pressure / cohesion /
viscosity / heat-conduction rates with a fused leapfrog, behind a C ABI
(include/pyticles_hip.h) loaded with ctypes.

Modules named after the reference keep its Python surfaces:
    neighbour_list, spkernel, properties, c_properties, forces, c_forces,
    spam_complete_force, particles, box
Device-resident handles: device.DeviceBox (one system), arena.Arena (many
systems, batched); serve.Server runs a job stream on an arena with HBM-budget
admission; autotune picks launch schedules.
"""
__version__ = "0.1.0"
