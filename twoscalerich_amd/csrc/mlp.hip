// mlp.hip - learned step controller (placeholder entry points; kernels follow).
#include "device.h"
namespace tsr {
struct Mlp { bool ready = false; };
Mlp* mlp_create() { return new Mlp(); }
void mlp_free(Mlp* m) { delete m; }
int mlp_set(Mlp* m, const float*, const float*, const float*) { m->ready = false; return 1; }
int mlp_infer(Mlp*, const float*, int, float*, hipStream_t) { return 1; }
bool mlp_ready(const Mlp* m) { return m && m->ready; }
}
