// The subset of the VapourSynth API 4 binary interface this plugin uses.
//
// A plugin receives the host's function tables as plain structs of function
// pointers, so the ABI is fixed by each table's slot order and by the layout
// of the structs passed by pointer. This header restates that layout (API 4,
// as vendored by the reference at third_party/vapoursynth/VapourSynth4.h):
// every slot of VSAPI up to logMessage is present in order; the slots this
// plugin does not call are typed as opaque pointers. Building against the
// official VapourSynth4.h instead gives the identical binary.
#pragma once

#include <stddef.h>
#include <stdint.h>

#define VS4_MAKE_VERSION(major, minor) (((major) << 16) | (minor))
#define VS4_API_VERSION VS4_MAKE_VERSION(4, 0)

extern "C" {

struct VSFrame;
struct VSNode;
struct VSCore;
struct VSPlugin;
struct VSMap;
struct VSFrameContext;
struct VSFunction;

enum { cfUndefined = 0, cfGray = 1, cfRGB = 2, cfYUV = 3 };
enum { stInteger = 0, stFloat = 1 };
enum { fmParallel = 0, fmParallelRequests = 1, fmUnordered = 2, fmFrameState = 3 };
enum { arInitial = 0, arAllFramesReady = 1, arError = -1 };
enum { mtDebug = 0, mtInformation = 1, mtWarning = 2, mtCritical = 3, mtFatal = 4 };
enum { rpGeneral = 0 };
enum { maReplace = 0, maAppend = 1 };

struct VSVideoFormat {
  int colorFamily;
  int sampleType;
  int bitsPerSample;
  int bytesPerSample;
  int subSamplingW;
  int subSamplingH;
  int numPlanes;
};

struct VSVideoInfo {
  VSVideoFormat format;
  int64_t fpsNum;
  int64_t fpsDen;
  int width;
  int height;
  int numFrames;
};

struct VSFilterDependency {
  VSNode* source;
  int requestPattern;
};

typedef void (*VSPublicFunction)(const VSMap* in, VSMap* out, void* userData, VSCore* core,
                                 const struct VSAPI* vsapi);
typedef const VSFrame* (*VSFilterGetFrame)(int n, int activationReason, void* instanceData,
                                           void** frameData, VSFrameContext* frameCtx,
                                           VSCore* core, const struct VSAPI* vsapi);
typedef void (*VSFilterFree)(void* instanceData, VSCore* core, const struct VSAPI* vsapi);
typedef void (*VS4Slot)(void);  // a slot this plugin never calls

struct VSPLUGINAPI {
  int (*getAPIVersion)(void);
  int (*configPlugin)(const char* identifier, const char* pluginNamespace, const char* name,
                      int pluginVersion, int apiVersion, int flags, VSPlugin* plugin);
  int (*registerFunction)(const char* name, const char* args, const char* returnType,
                          VSPublicFunction argsFunc, void* functionData, VSPlugin* plugin);
};

struct VSAPI {
  void (*createVideoFilter)(VSMap* out, const char* name, const VSVideoInfo* vi,
                            VSFilterGetFrame getFrame, VSFilterFree free, int filterMode,
                            const VSFilterDependency* dependencies, int numDeps,
                            void* instanceData, VSCore* core);        // 0
  VS4Slot createVideoFilter2, createAudioFilter, createAudioFilter2;  // 1-3
  VS4Slot setLinearFilter, setCacheMode, setCacheOptions;             // 4-6
  void (*freeNode)(VSNode* node);                                     // 7
  VS4Slot addNodeRef, getNodeType;                                    // 8-9
  const VSVideoInfo* (*getVideoInfo)(VSNode* node);                   // 10
  VS4Slot getAudioInfo;                                               // 11
  VSFrame* (*newVideoFrame)(const VSVideoFormat* format, int width, int height,
                            const VSFrame* propSrc, VSCore* core);    // 12
  VS4Slot newVideoFrame2, newAudioFrame, newAudioFrame2;              // 13-15
  void (*freeFrame)(const VSFrame* f);                                // 16
  VS4Slot addFrameRef, copyFrame, getFramePropertiesRO, getFramePropertiesRW;  // 17-20
  ptrdiff_t (*getStride)(const VSFrame* f, int plane);                // 21
  const uint8_t* (*getReadPtr)(const VSFrame* f, int plane);          // 22
  uint8_t* (*getWritePtr)(VSFrame* f, int plane);                     // 23
  VS4Slot getVideoFrameFormat, getAudioFrameFormat, getFrameType;     // 24-26
  int (*getFrameWidth)(const VSFrame* f, int plane);                  // 27
  int (*getFrameHeight)(const VSFrame* f, int plane);                 // 28
  VS4Slot getFrameLength, getVideoFormatName, getAudioFormatName;     // 29-31
  VS4Slot queryVideoFormat, queryAudioFormat, queryVideoFormatID;     // 32-34
  VS4Slot getVideoFormatByID, getFrame, getFrameAsync;                // 35-37
  const VSFrame* (*getFrameFilter)(int n, VSNode* node, VSFrameContext* frameCtx);  // 38
  void (*requestFrameFilter)(int n, VSNode* node, VSFrameContext* frameCtx);        // 39
  VS4Slot releaseFrameEarly, cacheFrame;                              // 40-41
  void (*setFilterError)(const char* errorMessage, VSFrameContext* frameCtx);       // 42
  VS4Slot createFunction, freeFunction, addFunctionRef, callFunction; // 43-46
  VS4Slot createMap, freeMap, clearMap, copyMap;                      // 47-50
  void (*mapSetError)(VSMap* map, const char* errorMessage);          // 51
  VS4Slot mapGetError, mapNumKeys, mapGetKey, mapDeleteKey;           // 52-55
  VS4Slot mapNumElements, mapGetType, mapSetEmpty;                    // 56-58
  int64_t (*mapGetInt)(const VSMap* map, const char* key, int index, int* error);   // 59
  VS4Slot mapGetIntSaturated, mapGetIntArray, mapSetInt, mapSetIntArray;            // 60-63
  double (*mapGetFloat)(const VSMap* map, const char* key, int index, int* error);  // 64
  VS4Slot mapGetFloatSaturated, mapGetFloatArray, mapSetFloat, mapSetFloatArray;    // 65-68
  VS4Slot mapGetData, mapGetDataSize, mapGetDataTypeHint, mapSetData; // 69-72
  VSNode* (*mapGetNode)(const VSMap* map, const char* key, int index, int* error);  // 73
  VS4Slot mapSetNode, mapConsumeNode, mapGetFrame, mapSetFrame;       // 74-77
  VS4Slot mapConsumeFrame, mapGetFunction, mapSetFunction, mapConsumeFunction;      // 78-81
  VS4Slot registerFunction, getPluginByID, getPluginByNamespace, getNextPlugin;     // 82-85
  VS4Slot getPluginName, getPluginID, getPluginNamespace, getNextPluginFunction;    // 86-89
  VS4Slot getPluginFunctionByName, getPluginFunctionName;             // 90-91
  VS4Slot getPluginFunctionArguments, getPluginFunctionReturnType;    // 92-93
  VS4Slot getPluginPath, getPluginVersion, invoke, createCore;        // 94-97
  VS4Slot freeCore, setMaxCacheSize, setThreadCount, getCoreInfo;     // 98-101
  VS4Slot getAPIVersion;                                              // 102
  void (*logMessage)(int msgType, const char* msg, VSCore* core);     // 103
};

}  // extern "C"

static_assert(offsetof(VSAPI, freeNode) == 7 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, getVideoInfo) == 10 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, newVideoFrame) == 12 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, freeFrame) == 16 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, getStride) == 21 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, getFrameWidth) == 27 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, getFrameFilter) == 38 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, setFilterError) == 42 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, mapSetError) == 51 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, mapGetInt) == 59 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, mapGetFloat) == 64 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, mapGetNode) == 73 * sizeof(void*), "VSAPI slot order");
static_assert(offsetof(VSAPI, logMessage) == 103 * sizeof(void*), "VSAPI slot order");
static_assert(sizeof(VSVideoInfo) == 64 && offsetof(VSVideoInfo, width) == 48, "VSVideoInfo layout");
