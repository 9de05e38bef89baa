// Command line of the reference (src/main.cpp:5-68) over the C ABI:
//   bingo_cli [d|n|p|s] -g <edge-list path> [-i|-d] [-r rounds] [-b batch] [-s seed]
// d/n/p/s select DeepWalk / node2vec / PPR / sampling; -i insert-only,
// -d delete-only, default mixed.  A bare dataset name is looked up under
// $BINGO_DATASET_DIR (default ../dataset, as src/apps.cu:12 opens it).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/bingo.h"

int main(int argc, char **argv)
{
    int app = BINGO_DEEPWALK, utype = BINGO_UPD_MIXED;
    std::string ds = "AM";
    bingo_instance_opts o;
    bingo_instance_opts_default(&o);
    int pos = 1;
    if (pos < argc && argv[pos][0] != '-') {
        switch (argv[pos][0]) {
        case 'n': app = BINGO_NODE2VEC; break;
        case 'p': app = BINGO_PPR; break;
        case 's': app = BINGO_SAMPLING; break;
        default: app = BINGO_DEEPWALK; break;
        }
        ++pos;
    }
    for (; pos < argc; ++pos) {
        if (argv[pos][0] != '-') continue;
        const char f = argv[pos][1];
        const char *val = pos + 1 < argc ? argv[pos + 1] : nullptr;
        if (f == 'g' && val) ds = argv[++pos];
        else if (f == 'i') utype = BINGO_UPD_INSERT;
        else if (f == 'd') utype = BINGO_UPD_DELETE;
        else if (f == 'r' && val) o.rounds = std::atoi(argv[++pos]);
        else if (f == 'b' && val) o.batch = std::atoll(argv[++pos]);
        else if (f == 's' && val) o.seed = std::strtoull(argv[++pos], nullptr, 10);
    }
    if (ds.find('/') == std::string::npos) {
        const char *dir = std::getenv("BINGO_DATASET_DIR");
        ds = std::string(dir ? dir : "../dataset") + "/" + ds;
    }
    bingo_instance_result r;
    const int rc = bingo_instance(app, utype, ds.c_str(), &o, &r);
    if (rc) {
        std::fprintf(stderr, "bingo_cli: error %d: %s\n", rc, bingo_last_error());
        return 1;
    }
    return 0;
}
