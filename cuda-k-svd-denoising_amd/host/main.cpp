// `denoising [config.json]`: the reference's entry point (src/main.cpp:56-124)
// with the same per-image and batch report, plus the K-SVD throughput.
#include <iostream>
#include <string>

#include "denoisingLib.h"

using namespace denoising;

static int testBatchDenoiser(const std::string &file)
{
    double init = 0, work = 0, fin = 0;
    BatchDenoiser *batch = BatchDenoiser::factory(file);
    if (!batch) return 1;
    std::vector<Denoiser *> list = batch->getDenoiserList();
    // files whose K-SVD can share dictionary-sweep launches run together
    // (their outputs are the bits each gets alone); the report below is
    // per file as before
    HipKSvdDenoiser::denoiseGroups(list);
    int failures = 0;
    for (size_t i = 0; i < list.size(); i++) {
        Denoiser *d = list[i];
        const signed char status = d->denoising();
        switch (status) {
        case 0: {
            baseUtl::TimeElapsed *t = d->getTimeElapsed();
            std::vector<double> *psnr = d->getPsnr();
            init += t->init;
            work += t->working;
            fin += t->finalize;
            std::cout << std::endl << "## Image: " << i + 1 << " execution time: " << t->getTotalTime() / 1000. << " s"
                      << std::endl;
            std::cout << "    init time: " << t->init << " ms" << std::endl;
            std::cout << "    working time: " << t->working / 1000. << " s" << std::endl;
            std::cout << "    finalize time: " << t->finalize << " ms" << std::endl;
            if ((*psnr)[0] >= 0) {
                std::cout << "    PSNR before:" << (*psnr)[0] << std::endl;
                std::cout << "    PSNR after:" << (*psnr)[1] << std::endl;
            }
            if (auto *h = dynamic_cast<HipKSvdDenoiser *>(d))
                std::cout << "    K-SVD patch-iters/s: " << h->patchItersPerSec() << " (" << h->patches()
                          << " patches" << (h->groupSize() > 1 ? ", sweeps shared by " + std::to_string(h->groupSize()) +
                                                                     " pictures" : std::string())
                          << ")" << std::endl;
            break;
        }
        case -1:
            std::cout << "Image Loading Failed: " << d->getError() << std::endl;
            failures++;
            break;
        case -2:
            std::cout << "Denoising Failed: " << d->getError() << std::endl;
            failures++;
            break;
        case -3:
            std::cout << "Image Saving Failed: " << d->getError() << std::endl;
            failures++;
            break;
        }
    }
    init /= 1000.;
    work /= 1000.;
    fin /= 1000.;
    std::cout << std::endl << "# Total Batch execution time: " << init + fin + work << " s" << std::endl;
    std::cout << "    Total Batch init time: " << init << " s" << std::endl;
    std::cout << "    Total Batch working time: " << work << " s" << std::endl;
    std::cout << "    Total Batch finalize time: " << fin << " s" << std::endl;
    delete batch;
    return failures ? 2 : 0;
}

int main(int argc, char *argv[])
{
    std::string file = "../config.json";
    if (argc == 2) file = argv[1];
    return testBatchDenoiser(file);
}
