// Runs the host OBJ reader (csrc/host/obj_loader.hpp) on each file named on
// the command line, for a build with AddressSanitizer and UndefinedBehavior-
// Sanitizer.  Prints "ok <vertices> <triangles>" or "error <message>" per
// file; a malformed file must end in an exception, never in a memory error.
#include <cstdio>
#include <exception>

#include "host/obj_loader.hpp"

int main(int argc, char** argv) {
    for (int k = 1; k < argc; ++k) {
        try {
            const crt_host::Mesh m = crt_host::load_obj(argv[k]);
            std::printf("ok %zu %zu\n", m.n_vertices(), m.n_tris());
        } catch (const std::exception& e) {
            std::printf("error %s\n", e.what());
        }
    }
    return 0;
}
