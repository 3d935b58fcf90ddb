// A C++ client of include/mclicks.hpp: prepares the dataset given on the
// command line, prints its sizes and train order, and reports whether a
// model could be created on device 0.
#include <cstdio>

#include "mclicks.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    mclicks::Dataset ds = mclicks::Dataset::parse(argv[1], -1, 0.2f);
    std::printf("parsed %lld train %lld test %lld n_qd %lld\n", (long long)ds.n_parsed(), (long long)ds.n_train(),
                (long long)ds.n_test(), (long long)ds.n_qd());
    mclicks::Sessions tr = ds.train();
    for (int64_t s = 0; s < tr.size(); s++) std::printf("train %d\n", tr.query[(size_t)s]);
    try {
        mclicks::ClickModel cm(mclicks::ALL == -1 ? MC_PBM : MC_DBN, 0);
        std::printf("model on device\n");
    } catch (mclicks::Error& e) {
        std::printf("model error %d\n", (int)e.status);
    }
    try {
        mclicks::Dataset::parse("/nonexistent/log.txt", -1, 0.2f);
    } catch (mclicks::Error& e) {
        std::printf("io error %d\n", (int)e.status);
    }
    return 0;
}
