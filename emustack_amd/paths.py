"""Directory prefix for mesh files (mirror of backend/paths.py:27-37).

NanoStruct.mesh_file names are resolved as msh_path + mesh_file, exactly as
the reference's Simmo.calc_modes does (mode_calcs.py:360-361).  Absolute mesh
file names work with the default empty prefix.
"""
import os

msh_path = os.environ.get('EMU_MSH_PATH', '')
