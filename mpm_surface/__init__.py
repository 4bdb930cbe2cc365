"""Surfacing of baked particle caches on the GPU (reference src/mpm_surface): particles -> density field ->
closed triangle meshes with vertex velocities, written next to the particle cache for the Blender add-on."""
