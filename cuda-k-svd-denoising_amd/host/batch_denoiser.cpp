// BatchDenoiser: config.json -> one Denoiser per file, same schema as the
// reference (src/lib/denoisingLib/batch_denoiser.cpp:36-117): the thirteen
// globalParams keys are required, every per-file key except "name" falls back
// to its global value, and the output is <outputFolder>/<pw>_<ph>_<name>.
// Optional keys of this backend (global or per file): "userCap" (1024),
// "geometry" ("true" | "compat"), "device" (0).
#include <fstream>
#include <iostream>

#include "denoisingLib.h"
#include "json.h"

using namespace denoising;

BatchDenoiser::BatchDenoiser() {}

BatchDenoiser::~BatchDenoiser()
{
    for (Denoiser *d : denoisers) delete d;
}

std::vector<Denoiser *> BatchDenoiser::getDenoiserList() { return denoisers; }

BatchDenoiser *BatchDenoiser::factory(std::string jsonFile)
{
    std::ifstream stream(jsonFile);
    if (!stream) {
        std::cerr << "config: cannot open " << jsonFile << std::endl;
        return nullptr;
    }
    json::Object config;
    if (!config.parse(stream)) {
        std::cerr << "config: " << jsonFile << ": " << config.last_error << std::endl;
        return nullptr;
    }
    DenoiserTypeMap map;
    BatchDenoiser *batch = new BatchDenoiser();
    try {
        const json::Object g = config.get<json::Object>("globalParams");
        const std::string inputFolder = config.get<json::String>("inputFolder");
        const std::string outputFolder = config.get<json::String>("outputFolder");
        const json::Array files = config.get<json::Array>("files");

        const int gPatchW = (int)g.get<json::Number>("patchWidthDim");
        const int gPatchH = (int)g.get<json::Number>("patchHeightDim");
        const int gSlideW = (int)g.get<json::Number>("slidingWidth");
        const int gSlideH = (int)g.get<json::Number>("slidingHeight");
        const bool gSpeckle = g.get<bool>("speckle");
        const bool gBW = g.get<bool>("B&W");
        const std::string gType = g.get<json::String>("type");
        const int gAtoms = (int)g.get<json::Number>("atoms");
        const int gIter = (int)g.get<json::Number>("ksvdIter");
        const int gOmpIter = (int)g.get<json::Number>("ompIter");
        const int gMinOmp = (int)g.get<json::Number>("minOmpIterBatch");
        const int gSubW = (int)g.get<json::Number>("subImageWidthDim");
        const int gSubH = (int)g.get<json::Number>("subImageHeightDim");
        const int gCap = (int)g.get<json::Number>("userCap", 1024);
        const std::string gGeom = g.get<json::String>("geometry", "true");
        const int gDevice = (int)g.get<json::Number>("device", 0);

        for (size_t i = 0; i < files.size(); i++) {
            const json::Object f = files.get<json::Object>(i);
            const int pw = (int)f.get<json::Number>("patchWidthDim", gPatchW);
            const int ph = (int)f.get<json::Number>("patchHeightDim", gPatchH);
            const std::string name = f.get<json::String>("name");
            const std::string ref = f.get<json::String>("ref", "");
            const std::string type = f.get<json::String>("type", gType);
            const std::string geom = f.get<json::String>("geometry", gGeom);
            const DenoiserType t = map.lookup(type);
            if (t == DENOISER_TYPE_UNKNOWN) throw json::Error("unknown denoiser type \"" + type + "\"");
            if (geom != "true" && geom != "compat") throw json::Error("geometry must be \"true\" or \"compat\"");
            Denoiser *d = Denoiser::factory(t, inputFolder + "/" + name,
                                            outputFolder + "/" + std::to_string(pw) + "_" + std::to_string(ph) +
                                                "_" + name);
            d->patchWidthDim = pw;
            d->patchHeightDim = ph;
            d->slidingWidth = (int)f.get<json::Number>("slidingWidth", gSlideW);
            d->slidingHeight = (int)f.get<json::Number>("slidingHeight", gSlideH);
            d->refImage = inputFolder + "/" + ref;
            d->ompIter = (int)f.get<json::Number>("ompIter", gOmpIter);
            d->atoms = (int)f.get<json::Number>("atoms", gAtoms);
            d->iter = (int)f.get<json::Number>("ksvdIter", gIter);
            d->speckle = f.get<bool>("speckle", gSpeckle);
            d->bw = f.get<bool>("B&W", gBW);
            d->minOmpIterBatch = (int)f.get<json::Number>("minOmpIterBatch", gMinOmp);
            d->subImageWidthDim = (int)f.get<json::Number>("subImageWidthDim", gSubW);
            d->subImageHeightDim = (int)f.get<json::Number>("subImageHeightDim", gSubH);
            d->userCap = (int)f.get<json::Number>("userCap", gCap);
            d->geometry = geom == "compat" ? 1 : 0;
            d->device = (int)f.get<json::Number>("device", gDevice);
            batch->denoisers.push_back(d);
        }
    } catch (const json::Error &e) {
        std::cerr << "config: " << jsonFile << ": " << e.what() << std::endl;
        delete batch;
        return nullptr;
    }
    return batch;
}
