"""Per-midpoint tasks over libhss_host.so (include/hss_host.h, hss_host_job_* and
hss_host_worker_*): the roles of the reference's SPITS module -- job manager, worker,
committer (common/src/execution/SpitzJobManager.cpp, SpitzWorker.cpp, SpitzCommitter.cpp)
-- for callers that distribute midpoints themselves (threads, processes, a cluster
scheduler).  A task record is one float32 midpoint; a result record is the midpoint, its
result arrays and its five statistics (float32).

    job = TaskJob(LINEAR_SEARCH, ["--input", path, ...])
    worker = job.worker(device=0)
    while (task := job.next_task()) is not None:
        job.commit_task(worker.run(task))
    out_dir = job.commit()
"""
import ctypes
import os

from . import HOST_LIB_PATH

LINEAR_SEARCH, DIFFERENTIAL_EVOLUTION, STRETCH_FREE = 0, 1, 2
PROGRAM_NAMES = {LINEAR_SEARCH: "single_host_linear_search",
                 DIFFERENTIAL_EVOLUTION: "single_host_de",
                 STRETCH_FREE: "single_host_stretch_free"}

_lib = None


class TaskError(RuntimeError):
    def __init__(self, rc, message):
        super().__init__(f"{message} (status {rc})")
        self.rc = rc


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise TaskError(-4, f"{HOST_LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(HOST_LIB_PATH)
        vp, i32, pi32 = ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)
        L.hss_host_last_error.restype = ctypes.c_char_p
        L.hss_host_job_open.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.POINTER(vp)]
        L.hss_host_job_record_sizes.argtypes = [vp, pi32, pi32, pi32]
        L.hss_host_job_next_task.argtypes = [vp, vp, i32, pi32]
        L.hss_host_job_commit_task.argtypes = [vp, vp, i32]
        L.hss_host_job_commit.argtypes = [vp, ctypes.c_char_p, i32]
        L.hss_host_job_close.argtypes = [vp]
        L.hss_host_worker_open.argtypes = [vp, i32, ctypes.POINTER(vp)]
        L.hss_host_worker_run.argtypes = [vp, vp, i32, vp, i32, pi32]
        L.hss_host_worker_close.argtypes = [vp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise TaskError(rc, lib().hss_host_last_error().decode(errors="replace"))


class TaskJob:
    """One command line of a program (program: LINEAR_SEARCH, DIFFERENTIAL_EVOLUTION or
    STRETCH_FREE; args: its options without the program name)."""

    def __init__(self, program, args):
        argv = [PROGRAM_NAMES.get(program, "program")] + [str(a) for a in args]
        self._argv = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
        h = ctypes.c_void_p()
        _check(lib().hss_host_job_open(int(program), len(argv), self._argv, ctypes.byref(h)))
        self._h = h
        t, r, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().hss_host_job_record_sizes(h, ctypes.byref(t), ctypes.byref(r), ctypes.byref(n)))
        self.task_bytes, self.result_bytes, self.tasks = t.value, r.value, n.value

    def next_task(self):
        """The next task record (bytes), or None when every CDP has been handed out."""
        buf = ctypes.create_string_buffer(self.task_bytes)
        n = ctypes.c_int32()
        _check(lib().hss_host_job_next_task(self._h, buf, self.task_bytes, ctypes.byref(n)))
        return buf.raw[:n.value] if n.value else None

    def commit_task(self, record):
        _check(lib().hss_host_job_commit_task(self._h, record, len(record)))

    def commit(self):
        """Writes the result files; returns the output directory."""
        buf = ctypes.create_string_buffer(4096)
        _check(lib().hss_host_job_commit(self._h, buf, len(buf)))
        return buf.value.decode()

    def worker(self, device=0):
        return TaskWorker(self, device)

    def close(self):
        if self._h:
            _check(lib().hss_host_job_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class TaskWorker:
    """A compute algorithm of the job's program on one device (close it before the job)."""

    def __init__(self, job, device=0):
        h = ctypes.c_void_p()
        _check(lib().hss_host_worker_open(job._h, int(device), ctypes.byref(h)))
        self._h, self._bytes, self._job = h, job.result_bytes, job

    def run(self, task):
        """The result record (bytes) of one task record."""
        buf = ctypes.create_string_buffer(self._bytes)
        n = ctypes.c_int32()
        _check(lib().hss_host_worker_run(self._h, task, len(task), buf, self._bytes,
                                         ctypes.byref(n)))
        return buf.raw[:n.value]

    def close(self):
        if self._h:
            _check(lib().hss_host_worker_close(self._h))
            self._h = None
