// Minimal JNI declarations written from the JNI specification (Java Native
// Interface Specification, chapter 4 "JNI Functions"), so the shim builds
// without a JDK.  JNIEnv is a pointer to a pointer to the function table;
// each function is called through its table index as numbered in the
// specification.  Only the functions the shim uses are wrapped.
#pragma once
#include <stdint.h>

extern "C" {

typedef int32_t jint;
typedef int64_t jlong;
typedef uint8_t jboolean;
typedef float jfloat;
typedef double jdouble;
typedef jint jsize;
typedef struct _jobject *jobject;
typedef jobject jclass;
typedef jobject jstring;
typedef jobject jarray;
typedef jarray jintArray;
typedef jarray jfloatArray;
typedef jarray jdoubleArray;
typedef jobject jthrowable;
typedef struct _jfieldID *jfieldID;

#define JNI_FALSE 0
#define JNI_TRUE 1
#define JNICALL
#define JNIEXPORT __attribute__((visibility("default")))

struct JNINativeInterface_ {
    void *fn[240];
};
typedef const struct JNINativeInterface_ *JNIEnv;

}  // extern "C"

namespace jni {

// Function-table indices (JNI specification, "Interface Function Table").
enum : int {
    kFindClass = 6,
    kThrowNew = 14,
    kExceptionCheck = 228,
    kGetObjectClass = 31,
    kGetFieldID = 94,
    kGetBooleanField = 96,
    kGetIntField = 100,
    kGetDoubleField = 103,
    kNewStringUTF = 167,
    kGetArrayLength = 171,
    kNewIntArray = 179,
    kGetIntArrayElements = 187,
    kGetFloatArrayElements = 189,
    kGetDoubleArrayElements = 190,
    kReleaseIntArrayElements = 195,
    kReleaseFloatArrayElements = 197,
    kReleaseDoubleArrayElements = 198,
    kSetIntArrayRegion = 211,
    kSetFloatArrayRegion = 213,
    kSetDoubleArrayRegion = 214,
    kGetDirectBufferAddress = 230,
};

template <class F>
inline F slot(JNIEnv *env, int i) {
    return (F)((*env)->fn[i]);
}

inline jclass FindClass(JNIEnv *e, const char *n) { return slot<jclass (*)(JNIEnv *, const char *)>(e, kFindClass)(e, n); }
inline jint ThrowNew(JNIEnv *e, jclass c, const char *m) {
    return slot<jint (*)(JNIEnv *, jclass, const char *)>(e, kThrowNew)(e, c, m);
}
inline jclass GetObjectClass(JNIEnv *e, jobject o) { return slot<jclass (*)(JNIEnv *, jobject)>(e, kGetObjectClass)(e, o); }
inline jfieldID GetFieldID(JNIEnv *e, jclass c, const char *n, const char *s) {
    return slot<jfieldID (*)(JNIEnv *, jclass, const char *, const char *)>(e, kGetFieldID)(e, c, n, s);
}
inline jboolean GetBooleanField(JNIEnv *e, jobject o, jfieldID f) {
    return slot<jboolean (*)(JNIEnv *, jobject, jfieldID)>(e, kGetBooleanField)(e, o, f);
}
inline jint GetIntField(JNIEnv *e, jobject o, jfieldID f) {
    return slot<jint (*)(JNIEnv *, jobject, jfieldID)>(e, kGetIntField)(e, o, f);
}
inline jdouble GetDoubleField(JNIEnv *e, jobject o, jfieldID f) {
    return slot<jdouble (*)(JNIEnv *, jobject, jfieldID)>(e, kGetDoubleField)(e, o, f);
}
inline jstring NewStringUTF(JNIEnv *e, const char *s) { return slot<jstring (*)(JNIEnv *, const char *)>(e, kNewStringUTF)(e, s); }
inline jsize GetArrayLength(JNIEnv *e, jarray a) { return slot<jsize (*)(JNIEnv *, jarray)>(e, kGetArrayLength)(e, a); }
inline jintArray NewIntArray(JNIEnv *e, jsize n) { return slot<jintArray (*)(JNIEnv *, jsize)>(e, kNewIntArray)(e, n); }
inline jint *GetIntArrayElements(JNIEnv *e, jintArray a) {
    return slot<jint *(*)(JNIEnv *, jintArray, jboolean *)>(e, kGetIntArrayElements)(e, a, nullptr);
}
inline jfloat *GetFloatArrayElements(JNIEnv *e, jfloatArray a) {
    return slot<jfloat *(*)(JNIEnv *, jfloatArray, jboolean *)>(e, kGetFloatArrayElements)(e, a, nullptr);
}
inline jdouble *GetDoubleArrayElements(JNIEnv *e, jdoubleArray a) {
    return slot<jdouble *(*)(JNIEnv *, jdoubleArray, jboolean *)>(e, kGetDoubleArrayElements)(e, a, nullptr);
}
inline void ReleaseIntArrayElements(JNIEnv *e, jintArray a, jint *p, jint mode) {
    slot<void (*)(JNIEnv *, jintArray, jint *, jint)>(e, kReleaseIntArrayElements)(e, a, p, mode);
}
inline void ReleaseFloatArrayElements(JNIEnv *e, jfloatArray a, jfloat *p, jint mode) {
    slot<void (*)(JNIEnv *, jfloatArray, jfloat *, jint)>(e, kReleaseFloatArrayElements)(e, a, p, mode);
}
inline void ReleaseDoubleArrayElements(JNIEnv *e, jdoubleArray a, jdouble *p, jint mode) {
    slot<void (*)(JNIEnv *, jdoubleArray, jdouble *, jint)>(e, kReleaseDoubleArrayElements)(e, a, p, mode);
}
inline void SetIntArrayRegion(JNIEnv *e, jintArray a, jsize st, jsize n, const jint *b) {
    slot<void (*)(JNIEnv *, jintArray, jsize, jsize, const jint *)>(e, kSetIntArrayRegion)(e, a, st, n, b);
}
inline void SetFloatArrayRegion(JNIEnv *e, jfloatArray a, jsize st, jsize n, const jfloat *b) {
    slot<void (*)(JNIEnv *, jfloatArray, jsize, jsize, const jfloat *)>(e, kSetFloatArrayRegion)(e, a, st, n, b);
}
inline void SetDoubleArrayRegion(JNIEnv *e, jdoubleArray a, jsize st, jsize n, const jdouble *b) {
    slot<void (*)(JNIEnv *, jdoubleArray, jsize, jsize, const jdouble *)>(e, kSetDoubleArrayRegion)(e, a, st, n, b);
}
inline void *GetDirectBufferAddress(JNIEnv *e, jobject b) {
    if (!b) return nullptr;
    return slot<void *(*)(JNIEnv *, jobject)>(e, kGetDirectBufferAddress)(e, b);
}

}  // namespace jni
