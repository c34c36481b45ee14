// ref_mesh_shim.cpp - C-linkage wrappers around the reference's own mesh-property
// functions (src/macro_mesh_properties.cpp:15, src/micro_mesh_properties.cpp:15),
// compiled from the reference tree into oracle/_ref/ (TEST INFRASTRUCTURE ONLY).
// Only declarations and the reference's struct headers are used here.
#include "macro_mesh.h"
#include "micro_mesh.h"

int macro_mesh_properties(const char* mesh_file, macro_mesh_struct& macro_mesh, int* macro_bc);
int micro_mesh_properties(const char* mesh_file, micro_mesh_struct& micro_mesh);

extern "C" int ref_macro_mesh_properties(const char* f, macro_mesh_struct* m, int* bc)
{
    return macro_mesh_properties(f, *m, bc);
}

extern "C" int ref_micro_mesh_properties(const char* f, micro_mesh_struct* m)
{
    return micro_mesh_properties(f, *m);
}
