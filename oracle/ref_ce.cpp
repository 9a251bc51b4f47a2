// ref_ce: runs the unmodified reference's rank-0 finisher, contig extension
// and contig writers on a given contigGraph.txt.
//
// Test infrastructure only (built into oracle/_ref/ by `make -C oracle ref`,
// never part of the product).  graph.cpp is compiled inside this file with its
// `main` renamed; kmerGraph.cpp / sequence.cpp / mympi.cpp are compiled beside
// it, all from where they lie under /root/reference.  This file only drives
// them:
//
//   ref_ce <contigGraph.txt> <S> -k K -c C -i <reads.fa> -o <new output dir> [-j]
//
// It performs what the reference's main does on rank 0 after its gather
// (graph.cpp:3612-3682): the nodes table is cleared (clear() keeps its bucket
// count) and refilled by recvBufDistNodeGraph from the contigGraph.txt bytes
// (the gathered buffer the reference writes to that file).  The bucket count
// the table had is recreated the way it arose: S inserts through operator[]
// (buildKmoleculeGraph's S survivors, graph.cpp:1123-1133) and one more
// operator[] lookup (the merge's first node access), then clear().  Then
// masterSimplifyNodeGraph, printContigs (noCEcontigs.fasta), four rounds of
// tips / low-coverage / loop-bubble / multiple-edge-bubble removal with the
// finisher between them, masterRemoveCrossEdge and masterPrintContigs
// (CEContig.fasta), and printJungGraph with -j.  None of these steps sends a
// message, so they run in one process; the merge that produces contigGraph.txt
// is the part that cannot (see the reference run notes of the design).

#define main swap_reference_main
#include "graph.cpp"
#undef main

#include <vector>

int main(int argc, char *argv[]) {
    if (argc < 3) {
        fprintf(stderr, "usage: ref_ce <contigGraph.txt> <S> -k K -c C -i reads.fa -o outdir [-j]\n");
        return 2;
    }
    const unsigned long long S = strtoull(argv[2], nullptr, 10);
    MPIEnviroment MPIcontrol;
    MPIcontrol.init(argc, argv);
    std::vector<char *> args;
    args.push_back(argv[0]);
    for (int i = 3; i < argc; i++) args.push_back(argv[i]);
    parameter parameters;
    parameters.getParameters((int)args.size(), args.data(), &MPIcontrol);

    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *rBuf = new char[n + 1];
    const size_t got = fread(rBuf, 1, (size_t)n, f);
    fclose(f);
    rBuf[got] = 0;

    distNodeGraph *nodeGraph = new distNodeGraph(&parameters, &MPIcontrol);
    for (unsigned long long i = 0; i < S; i++) nodeGraph->nodes[i];
    if (S) nodeGraph->nodes[0];
    nodeGraph->nodes.clear();
    nodeGraph->recvBufDistNodeGraph(rBuf, &parameters, &MPIcontrol);  // frees rBuf
    nodeGraph->masterSimplifyNodeGraph(&parameters, &MPIcontrol);
    nodeGraph->printContigs(&parameters, &MPIcontrol);
    for (int i = 0; i < 4; i++) {
        nodeGraph->masterTipsRemoval(&parameters, &MPIcontrol);
        nodeGraph->masterSimplifyNodeGraph(&parameters, &MPIcontrol);
        nodeGraph->masterLowCoverageEdgeRemoval(&parameters, &MPIcontrol);
        nodeGraph->masterSimplifyNodeGraph(&parameters, &MPIcontrol);
        nodeGraph->masterLoopBubbleRemoval(&parameters, &MPIcontrol);
        nodeGraph->masterMultipleEdgeBubbleRemoval(&parameters, &MPIcontrol);
        nodeGraph->masterSimplifyNodeGraph(&parameters, &MPIcontrol);
        nodeGraph->masterGraphStatistic(&parameters, &MPIcontrol);
    }
    nodeGraph->masterRemoveCrossEdge(&parameters, &MPIcontrol);
    nodeGraph->masterPrintContigs(&parameters, &MPIcontrol);
    if (parameters.JungGraphFlag) nodeGraph->printJungGraph(&parameters, &MPIcontrol);
    nodeGraph->nodes.clear();
    MPIcontrol.finalize();
    return 0;
}
