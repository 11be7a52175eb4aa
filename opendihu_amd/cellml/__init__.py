"""CellML (OpenCOR C export) front end: parser and HIP code generator.

The reference loads cell models as OpenCOR-generated C files and turns them
into compiled kernels at run time (core/src/cellml/source_code_generator/
00_parse_source_code_file.cpp:159-357, 03_generator_vc.cpp:13-302,873-1197).
This package does the same ahead of time for the HIP fibre kernels.
"""
from .opencor import OpenCorModel, parse_opencor_c
from .codegen import generate_hip_model
from .vc_kernel import convert_vc_kernel

__all__ = ["OpenCorModel", "parse_opencor_c", "generate_hip_model", "convert_vc_kernel"]
