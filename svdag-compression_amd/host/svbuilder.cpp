// svbuilder -- drop-in for the reference's svbuilder (src/svbuilder/main.cpp):
// same argv, same output files and names, same result block and stats.txt.
//
//   svbuilder input_model.{obj,ply} numLevels numBuildSteps [-l [infl diff refs]] [-c] [outputs...]
//
// numBuildSteps == 0: one SVO voxelized root-down (buildSVO) then toDAG;
// numBuildSteps  > 0: the reference's subtree builder (buildDAG).  The
// reductions and all encoders run on the GPU.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "symvox_amd.hpp"

static void printUsage() {
  printf(
      "\n"
      "Usage:\n"
      "      svbuilder input_model.obj numLevels numBuildSteps [--lossy or -l] [--cross-level-merging or -c] "
      "[<output.[svo | svdag | ussvdag | ssvdag]>\n"
      "Where:\n"
      "      input_model.obj: a 3D model in ASCII OBJ format (or ASCII PLY)\n"
      "      numLevels: levels of the octree to build (i.e. 10->1K^3, 13->8K^3...)\n"
      "      numBuildSteps:\n"
      "          if   0, raw build of the full octree\n"
      "          if > 0, levels of the 'base octree' to build. Then the subtrees\n"
      "                  of the full children are voxelized in parallel\n"
      "      <output.[svo | svdag | ussvdag | ssvdag]>: (optional) output files to be written\n"
      "\n");
}

static std::string dir_of(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? "." : p.substr(0, k);
}
static std::string base_without_ext(const std::string& p) {
  size_t k = p.find_last_of('/');
  std::string b = k == std::string::npos ? p : p.substr(k + 1);
  size_t d = b.find_last_of('.');
  return d == std::string::npos ? b : b.substr(0, d);
}

static std::string hq(size_t n) {
  char b[64];
  if (n >= 1000000000ull) snprintf(b, sizeof b, "%.2fG", n / 1e9);
  else if (n >= 1000000ull) snprintf(b, sizeof b, "%.2fM", n / 1e6);
  else if (n >= 1000ull) snprintf(b, sizeof b, "%.2fK", n / 1e3);
  else snprintf(b, sizeof b, "%zu", n);
  return b;
}
// human-readable duration of a time in seconds
static std::string hd(double sec) {
  char b[64];
  if (sec >= 3600.0) snprintf(b, sizeof b, "%.2f h", sec / 3600.0);
  else if (sec >= 60.0) snprintf(b, sizeof b, "%.2f min", sec / 60.0);
  else if (sec >= 1.0) snprintf(b, sizeof b, "%.2f s", sec);
  else if (sec >= 1e-3) snprintf(b, sizeof b, "%.2f ms", sec * 1e3);
  else snprintf(b, sizeof b, "%.2f us", sec * 1e6);
  return b;
}
static std::string hs(size_t n) {
  char b[64];
  if (n >= (1ull << 30)) snprintf(b, sizeof b, "%.2f GB", n / double(1ull << 30));
  else if (n >= (1ull << 20)) snprintf(b, sizeof b, "%.2f MB", n / double(1ull << 20));
  else if (n >= (1ull << 10)) snprintf(b, sizeof b, "%.2f KB", n / double(1ull << 10));
  else snprintf(b, sizeof b, "%zu B", n);
  return b;
}

int main(int argc, char** argv) {
  printf(
      "\n===============================================================================\n"
      "========   svbuilder (MI355X sparse-voxel-DAG engine, SymVox-compatible)   ====\n"
      "===============================================================================\n");
  if (argc < 4) {
    printUsage();
    exit(1);
  }
  std::string inputFile(argv[1]);
  int nLevels = atoi(argv[2]);
  int levelStep = atoi(argv[3]);
  printf(" MODEL: '%s'   [ %d levels, step %i ]   (%.0fK^3)\n", inputFile.c_str(), nLevels, levelStep,
         pow(2, nLevels) / 1024.f);
  printf("===============================================================================\n\n");
  const auto t0 = std::chrono::steady_clock::now();

  Scene scene;
  GeomOctree octree(&scene);
  bool isInputDAG = false;
  if (strstr(inputFile.c_str(), ".obj") || strstr(inputFile.c_str(), ".OBJ")) {
    scene.loadObj(inputFile, true, false, false, false, true);
  } else if (strstr(inputFile.c_str(), ".ply") || strstr(inputFile.c_str(), ".PLY")) {
    if (!scene.loadPly(inputFile)) exit(1);
  } else if (strstr(inputFile.c_str(), ".svdag")) {   // main.cpp:97-101
    EncodedSVDAG inputSVDAG;
    if (!inputSVDAG.load(inputFile)) exit(1);
    octree = inputSVDAG.decode(scene);
    if (octree.getState() != GeomOctree::S_DAG) exit(1);
    isInputDAG = true;
  } else {
    printf("Can't read input file '%s'. Only supported ASCII Obj (or PLY) files.\n", inputFile.c_str());
    exit(1);
  }

  bool lossy = false, multiLevel = false;
  float lossyInflation = 2.0f, allowedLossyDiffFactor = 1.0f;
  int includedNodeRefCount = 1;
  for (int i = 4; i < argc; ++i) {
    std::string arg = argv[i];
    if (arg == "--lossy" || arg == "-l") {
      try {
        if (i + 1 < argc) lossyInflation = std::min(10.0f, std::max(1.1f, std::stof(argv[i + 1])));
        if (i + 2 < argc) allowedLossyDiffFactor = std::min(32.0f, std::max(0.01f, std::stof(argv[i + 2])));
        if (i + 3 < argc) includedNodeRefCount = std::min(1000000, std::max(1, std::stoi(argv[i + 3])));
      } catch (std::invalid_argument&) {
        printf(" --- WARNING: Could not parse lossy compression parameters, using defaults. ---\n");
      }
      printf("Lossy parameters: inflation %.1f, allowed diff factor: %.2f, included node ref count: %i\n",
             lossyInflation, allowedLossyDiffFactor, includedNodeRefCount);
      lossy = true;
    } else if (arg == "--cross-level-merging" || arg == "-c") {
      multiLevel = true;
    }
  }
  printf("Lossy: %d, Cross-level: %d, Hidden geom: %d\n", lossy, multiLevel, 0);

  const svd::aabox3f ab = scene.getAABB();
  svd::aabox3d sceneBBoxD(svd::point3d(ab[0][0], ab[0][1], ab[0][2]), svd::point3d(ab[1][0], ab[1][1], ab[1][2]));
  if (!isInputDAG) {   // main.cpp:147-202
    if (levelStep == 0) {
      octree.buildSVO(nLevels, sceneBBoxD, false);
      octree.toDAG();
    } else {
      octree.buildDAG(nLevels, levelStep, sceneBBoxD, true);
    }
  }
  // main.cpp:177-193: normalize very large / very small boxes
  if (!isInputDAG) {
    const float dx = ab[1][0] - ab[0][0], dy = ab[1][1] - ab[0][1], dz = ab[1][2] - ab[0][2];
    const float bboxSize = std::sqrt(dx * dx + dy * dy + dz * dz);
    const float maxBboxSize = 100000.f;
    if (bboxSize > maxBboxSize || bboxSize < 0.1) {
      printf("Normalizing bbox; too small/large: %f\n", bboxSize);
      const float inv = maxBboxSize / bboxSize;
      svd::aabox3f nb(svd::point3f(0, 0, 0), svd::point3f(dx * inv, dy * inv, dz * inv));
      scene.setAABB(nb);
      octree.resizeSceneBbox(nb);
    }
  }
  octree.initChildLevels();

  const std::string basePath = dir_of(inputFile) + "/" + base_without_ext(inputFile) + "_" + std::to_string(nLevels);
  // main.cpp:213-217: "-l_<infl>_<factor>_<refs>" with one decimal kept
  auto trim1 = [](float v) { std::string t = std::to_string(v); return t.substr(0, t.find(".") + 2); };
  const std::string paramStr = "-l_" + trim1(lossyInflation) + "_" + trim1(allowedLossyDiffFactor) + "_" +
                               std::to_string(includedNodeRefCount);
  EncodedSVDAG svdag2;
  svdag2.encode(octree);
  svdag2.save(basePath + ".svdag");

  if (lossy) {                         // main.cpp:225-229
    EncodedSSVDAG esvdag2;
    esvdag2.encode(octree);
    esvdag2.save(basePath + ".esvdag");
  }
  if (lossy) {                         // main.cpp:231-238
    const double origDagTime = octree.getStats().toDAGTime;
    octree.toLossyDag(lossyInflation, allowedLossyDiffFactor, includedNodeRefCount);
    octree.getStats().toDAGTime = origDagTime;
  } else if (multiLevel) {
    octree.mergeAcrossAllLevels();
  }

  EncodedSVDAG svdag;
  svdag.encode(octree);
  EncodedSSVDAG esvdag;
  if (!multiLevel) esvdag.encode(octree);
  if (!multiLevel) octree.toSDAG(false, false);
  EncodedUSSVDAG ussvdag;
  if (!multiLevel) ussvdag.encode(octree);
  EncodedSSVDAG ssvdag;
  if (!multiLevel) ssvdag.encode(octree);

  std::string infix = lossy ? paramStr : (multiLevel ? "-multi" : "");
  svdag.save(basePath + infix + ".svdag");
  if (!multiLevel) {
    ussvdag.save(basePath + infix + ".ussvdag");
    ssvdag.save(basePath + infix + ".ssvdag");
    esvdag.save(basePath + infix + ".esvdag");
  }
  for (int i = 4; i < argc; ++i) {
    std::string out = argv[i];
    if (strstr(out.c_str(), ".svo") || strstr(out.c_str(), ".SVO")) continue;
    if (strstr(out.c_str(), ".ussvdag") || strstr(out.c_str(), ".USSVDAG")) ussvdag.save(out);
    else if (strstr(out.c_str(), ".ssvdag") || strstr(out.c_str(), ".SSVDAG")) ssvdag.save(out);
    else if (strstr(out.c_str(), ".esvdag") || strstr(out.c_str(), ".ESVDAG")) esvdag.save(out);
    else if (strstr(out.c_str(), ".svdag") || strstr(out.c_str(), ".SVDAG")) svdag.save(out);
  }

  const double totalTime = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  GeomOctree::Stats stats = octree.getStats();
  const size_t nv = octree.getNVoxels();
  printf("\n========= RESULTS '%s' [%d levels] (%.0fK^3) =========\n", inputFile.c_str(), nLevels, pow(2, nLevels) / 1024.f);
  printf("Voxels:     %s\t(%zu)\n", hq(stats.nTotalVoxels).c_str(), stats.nTotalVoxels);
  printf("SVO Nodes:  %s\t(%zu)\n", hq(stats.nNodesSVO).c_str(), stats.nNodesSVO);
  printf("DAG Nodes:  %s\t(%zu)\n", hq(stats.nNodesDAG).c_str(), stats.nNodesDAG);
  printf("SDAG Nodes: %s\t(%zu)\n", hq(stats.nNodesSDAG).c_str(), stats.nNodesSDAG);
  printf("Pointerless SVO    : %s\t(%.3f bits/vox)\n", hs(stats.nNodesSVO).c_str(), (8 * stats.nNodesSVO) / float(nv));
  printf("Encoded SVDAG      : %s\t(%.3f bits/vox)\n", hs(svdag.getDataSize()).c_str(), (8 * svdag.getDataSize()) / float(nv));
  printf("Encoded ESVDAG     : %s\t(%.3f bits/vox)\n", hs(esvdag.getDataSize()).c_str(), (8 * esvdag.getDataSize()) / float(nv));
  printf("Encoded USSVDAG    : %s\t(%.3f bits/vox)\n", hs(ussvdag.getDataSize()).c_str(), (8 * ussvdag.getDataSize()) / float(nv));
  printf("Encoded SSVDAG     : %s\t(%.3f bits/vox)\n", hs(ssvdag.getDataSize()).c_str(), (8 * ssvdag.getDataSize()) / float(nv));
  printf("SSVDAG / DAG       : %.1f %%\n", 100.f * ssvdag.getDataSize() / (float)svdag.getDataSize());
  printf("SSVDAG / SVO       : %.1f %%\n", 100.f * ssvdag.getDataSize() / (float)stats.nNodesSVO);
  printf("SVO->SVDAG time    : %s\n", hd(stats.toDAGTime).c_str());
  printf("SVDAG->LSVDAG time : %s\n", hd(stats.toLSVDAGTime).c_str());
  printf("SVDAG->SSVDAG time : %s\n", hd(stats.toSDAGTime).c_str());
  printf("Total time         : %s\n", hd(totalTime).c_str());
  printf("===================================================================\n\n");

  // main.cpp:320-366: the same block to the console, then appended to ./stats.txt
  const std::string baseName = base_without_ext(inputFile);
  auto ms = [](double sec) { return (size_t)(sec * 1e3); };
  for (int pass = 0; pass < 2; ++pass) {
    FILE* st = pass == 0 ? stdout : fopen("stats.txt", "a");
    if (!st) break;
    fprintf(st, "%s, %d, lossy: %d, cross-level: %d\n", baseName.c_str(), nLevels, lossy, multiLevel);
    fprintf(st, "#Voxels, %zu\n", stats.nTotalVoxels);
    fprintf(st, ", SVDAG, ESVDAG, SSVDAG, SVO\n");
    fprintf(st, "#nodes, %zu, '', %zu, %zu\n", stats.nNodesDAG, stats.nNodesSDAG, stats.nNodesSVO);
    fprintf(st, "memory (bytes), %zu, %zu, %zu, %zu\n", svdag.getDataSize(), esvdag.getDataSize(),
            ssvdag.getDataSize(), stats.nNodesSVO);
    if (lossy) {
      fprintf(st, "Construction times:\n");
      fprintf(st, ",SVDAG, LSVDAG, Total, Hashing, SimNodes, Clustering\n");
      fprintf(st, "time (ms), %zu, %zu, %zu, %zu, %zu, %zu\n", ms(stats.toDAGTime), ms(stats.toLSVDAGTime),
              ms(totalTime), ms(stats.lHashing), ms(stats.lSimNodes), ms(stats.lClustering));
      fprintf(st, "Lossy stats, %.2f, %.2f, %i\n", lossyInflation, allowedLossyDiffFactor, includedNodeRefCount);
      fprintf(st, "TotalVoxDifference, #NodesIn, #ClustersOut, #edges, svdag mem, lsvdag mem\n");
      fprintf(st, "%zu, %zu, %zu, %zu, %zu, %zu\n", stats.totalLossyVoxelDifference, stats.nClusteredNodes,
              stats.nClusters, stats.nEdges, svdag2.getDataSize(), svdag.getDataSize());
    } else if (multiLevel) {
      fprintf(st, "Construction times:\n");
      fprintf(st, ",SVDAG, Total, Cross-level\n");
      fprintf(st, "time (ms), %zu, %zu, %zu\n", ms(stats.toDAGTime), ms(totalTime), ms(stats.crossMergeTime));
      fprintf(st, "Cross-level, nodes eliminated, svdag mem, csvdag mem\n");
      fprintf(st, ", %zu, %zu, %zu\n", stats.nCrossLevelMerged, svdag2.getDataSize(), svdag.getDataSize());
    } else {
      fprintf(st, "Construction times:\n");
      fprintf(st, ",SVDAG, Total\n");
      fprintf(st, "time (ms), %zu, %zu\n", ms(stats.toDAGTime), ms(totalTime));
    }
    fprintf(st, "\n\n");
    if (pass == 1) fclose(st);
  }
  return 0;
}
