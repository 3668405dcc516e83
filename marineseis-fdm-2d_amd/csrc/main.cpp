// bin/main_debug <model.json> <out_root>: src/main.cu:58-77 on libmsfdm.
#include <cstdio>

#include "msfdm.h"

int main(int argc, char *argv[]) {
    if (argc != 3) {
        std::fprintf(stderr,
                     "usage: %s <model.json> <output_dir>\n"
                     "example: %s models/case/case.json output/case\n",
                     argv[0], argv[0]);
        return 1;
    }
    const int rc = msfdm_run_cli(argv[1], argv[2]);
    if (rc != 0) std::fprintf(stderr, "%s\n", msfdm_last_error());
    return rc;
}
