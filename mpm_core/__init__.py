"""Blender-independent MPM core: interface, formats, runner; "hip" is the MI355X backend."""
