"""A fake JNIEnv for driving the Java_gpufitImFCS_GpufitImFCS_* exports
without a JVM.

The function table is laid out by the slot numbers of the JNI specification;
only the functions the shim uses are implemented, every other slot traps.
Java objects are integer handles into a registry: numpy arrays stand for
Java primitive arrays and direct NIO buffers, Python objects with attributes
for GpufitImFCS.ACFParameters.
"""
import ctypes

import numpy as np

from imaging_fcs_1_52_amd import native

VP = ctypes.c_void_p

# JNI specification function-table slot numbers
SLOTS = {"FindClass": 6, "ThrowNew": 14, "GetObjectClass": 31, "GetFieldID": 94, "GetBooleanField": 96,
         "GetIntField": 100, "GetDoubleField": 103, "NewStringUTF": 167, "GetArrayLength": 171,
         "NewIntArray": 179, "GetIntArrayElements": 187, "GetFloatArrayElements": 189,
         "GetDoubleArrayElements": 190, "ReleaseIntArrayElements": 195, "ReleaseFloatArrayElements": 197,
         "ReleaseDoubleArrayElements": 198, "SetIntArrayRegion": 211, "SetFloatArrayRegion": 213,
         "SetDoubleArrayRegion": 214, "GetDirectBufferAddress": 230}

ACF_FIELDS = {"width": "I", "height": "I", "win_star": "I", "hin_star": "I", "w_temp": "I", "h_temp": "I",
              "pixbinX": "I", "pixbinY": "I", "binningX": "I", "binningY": "I", "firstframe": "I",
              "lastframe": "I", "framediff": "I", "cfXDistance": "I", "cfYDistance": "I", "correlatorp": "D",
              "correlatorq": "D", "frametime": "D", "background": "I", "mtab1": "D", "mtabchanumminus1": "D",
              "sampchanumminus1": "D", "chanum": "I", "isNBcalculation": "Z", "bleachcorr_gpu": "Z",
              "bleachcorr_order": "I", "nopit": "I", "ave": "I"}


class JavaObject:
    def __init__(self, cls, **fields):
        self.cls = cls
        self.__dict__.update(fields)


class FakeEnv:
    def __init__(self):
        self.objects = {}
        self.next = 0x1000
        self.exceptions = []
        self.field_requests = []
        self.trapped = []
        self._cb = []
        self.table = (VP * 240)()
        for i in range(240):
            self.table[i] = self._keep(ctypes.CFUNCTYPE(None)(self._trap(i)))
        self._install()
        self.cell = ctypes.c_void_p(ctypes.addressof(self.table))
        self.env = ctypes.c_void_p(ctypes.addressof(self.cell))
        self.lib = native.lib()

    # -- registry
    def new(self, obj):
        h = self.next
        self.next += 16
        self.objects[h] = obj
        return h

    def get(self, h):
        return self.objects.get(h)

    def _keep(self, f):
        self._cb.append(f)
        return ctypes.cast(f, VP).value

    def _trap(self, i):
        def t():
            self.trapped.append(i)
        return t

    # -- JNI functions
    def _install(self):
        F = ctypes.CFUNCTYPE
        arr_types = {np.dtype(np.int32): "I", np.dtype(np.float32): "F", np.dtype(np.float64): "D"}

        def find_class(env, name):
            return self.new(("class", name.decode()))

        def throw_new(env, cls, msg):
            self.exceptions.append((self.get(cls)[1], msg.decode()))
            return 0

        def get_object_class(env, obj):
            return self.new(("class", self.get(obj).cls))

        def get_field_id(env, cls, name, sig):
            self.field_requests.append((name.decode(), sig.decode()))
            return self.new(("field", name.decode(), sig.decode()))

        def field_getter(expect):
            def g(env, obj, fid):
                _, name, sig = self.get(fid)
                assert sig == expect, (name, sig, expect)
                return getattr(self.get(obj), name)
            return g

        def new_string_utf(env, s):
            return self.new(s.decode())

        def get_array_length(env, a):
            return len(self.get(a))

        def new_int_array(env, n):
            return self.new(np.zeros(n, np.int32))

        def get_elements(env, a, is_copy):
            return self.get(a).ctypes.data

        def release(env, a, p, mode):
            pass

        def set_region(dtype):
            def s(env, a, start, n, buf):
                arr = self.get(a)
                assert arr.dtype == dtype
                src = np.ctypeslib.as_array(ctypes.cast(buf, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                            shape=(n,))
                arr[start:start + n] = src
            return s

        def direct_address(env, b):
            return self.get(b).ctypes.data

        impl = {
            "FindClass": F(VP, VP, ctypes.c_char_p)(find_class),
            "ThrowNew": F(ctypes.c_int32, VP, VP, ctypes.c_char_p)(throw_new),
            "GetObjectClass": F(VP, VP, VP)(get_object_class),
            "GetFieldID": F(VP, VP, VP, ctypes.c_char_p, ctypes.c_char_p)(get_field_id),
            "GetBooleanField": F(ctypes.c_uint8, VP, VP, VP)(field_getter("Z")),
            "GetIntField": F(ctypes.c_int32, VP, VP, VP)(field_getter("I")),
            "GetDoubleField": F(ctypes.c_double, VP, VP, VP)(field_getter("D")),
            "NewStringUTF": F(VP, VP, ctypes.c_char_p)(new_string_utf),
            "GetArrayLength": F(ctypes.c_int32, VP, VP)(get_array_length),
            "NewIntArray": F(VP, VP, ctypes.c_int32)(new_int_array),
            "GetIntArrayElements": F(VP, VP, VP, VP)(get_elements),
            "GetFloatArrayElements": F(VP, VP, VP, VP)(get_elements),
            "GetDoubleArrayElements": F(VP, VP, VP, VP)(get_elements),
            "ReleaseIntArrayElements": F(None, VP, VP, VP, ctypes.c_int32)(release),
            "ReleaseFloatArrayElements": F(None, VP, VP, VP, ctypes.c_int32)(release),
            "ReleaseDoubleArrayElements": F(None, VP, VP, VP, ctypes.c_int32)(release),
            "SetIntArrayRegion": F(None, VP, VP, ctypes.c_int32, ctypes.c_int32, VP)(set_region(np.int32)),
            "SetFloatArrayRegion": F(None, VP, VP, ctypes.c_int32, ctypes.c_int32, VP)(set_region(np.float32)),
            "SetDoubleArrayRegion": F(None, VP, VP, ctypes.c_int32, ctypes.c_int32, VP)(set_region(np.float64)),
            "GetDirectBufferAddress": F(VP, VP, VP)(direct_address),
        }
        for name, f in impl.items():
            self.table[SLOTS[name]] = self._keep(f)

    # -- calling the exports
    def call(self, method, *args, restype=None):
        f = getattr(self.lib, "Java_gpufitImFCS_GpufitImFCS_" + method)
        f.restype = restype
        conv = []
        argtypes = [VP, VP]
        for a in args:
            if isinstance(a, float):
                argtypes.append(ctypes.c_float)
                conv.append(a)
            elif isinstance(a, (int, np.integer)) and not isinstance(a, bool):
                argtypes.append(ctypes.c_int32)
                conv.append(int(a))
            else:
                argtypes.append(VP)
                conv.append(a)
        f.argtypes = argtypes
        cls = self.new(("class", "gpufitImFCS/GpufitImFCS"))
        return f(self.env, cls, *conv)

    def handle(self, obj):
        """Registers a numpy array or JavaObject and returns its handle
        (wrapped so that `call` passes it as a pointer)."""
        return ctypes.c_void_p(self.new(obj))

    def acf_parameters(self, **fields):
        missing = set(ACF_FIELDS) - set(fields)
        assert not missing, missing
        return self.handle(JavaObject("gpufitImFCS/GpufitImFCS$ACFParameters", **fields))
