"""Socket front end of the serving loop: k-NN requests over TCP.

    python -m cudaruntime1_amd.netserve --db db.tsdb [--host 127.0.0.1] [--port 7070]

loads a tsdb database file onto the GPU, starts a continuous-batching server
over it (ts_server_*) and answers requests from any number of connections;
`Client` is the matching client. The reference has no serving path at all
(one host call per batch, main.cu:67-230).

Frames are little-endian and start with a magic word and the frame's length
in bytes, so a reader can always find the next one.

    request   'TSQ1' u32 length u64 id
              i32 measure i32 mode i32 k i32 kc i32 r i32 band i32 delta
              f32 epsilon i32 n_points, then n_points (x, y) f32 pairs
              (kc, r, band, delta = -2^31 and epsilon = NaN: the library's
              default for the measure and mode)
    response  'TSR1' u32 length u64 id i32 status i32 k f64 latency_ms,
              then k i32 indices and k f32 values; a non-zero status (a
              ts_status code, e.g. 6 = busy: admission refused) carries no
              results but a UTF-8 message

Requests on one connection are pipelined: each is submitted as it arrives
and answered when it completes, possibly out of order (match by id). A frame
that cannot be parsed closes its connection; others are unaffected.
"""
import argparse
import asyncio
import socket
import struct
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

REQ_MAGIC, RESP_MAGIC = b"TSQ1", b"TSR1"
REQ_HEAD = struct.Struct("<4sIQiiiiiiifi")   # magic, length, id, measure, mode, k, kc, r, band, delta, eps, n
RESP_HEAD = struct.Struct("<4sIQiid")        # magic, length, id, status, k, latency_ms
MAX_POINTS = 1 << 24
MAX_K = 1000
DEFAULT = -(1 << 31)  # an integer field left to the library's default


def encode_request(rid, points, measure, mode, k, kc, r, band, delta, epsilon):
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
    body = pts.tobytes()
    return REQ_HEAD.pack(REQ_MAGIC, REQ_HEAD.size + len(body), rid, measure, mode, k, kc, r, band, delta,
                         float(epsilon), len(pts)) + body


def decode_request(head, body):
    magic, length, rid, measure, mode, k, kc, r, band, delta, eps, n = REQ_HEAD.unpack(head)
    if magic != REQ_MAGIC or n < 1 or n > MAX_POINTS or length != REQ_HEAD.size + 8 * n or not 1 <= k <= MAX_K:
        raise ValueError("bad request frame")
    pts = np.frombuffer(body, np.float32).reshape(n, 2)
    return rid, dict(measure=measure, mode=mode, k=k, kc=kc, r=r, band=band, delta=delta, epsilon=eps), pts


def encode_response(rid, status, idx=None, val=None, latency_ms=0.0, message=""):
    if status == 0:
        k = len(idx)
        body = np.asarray(idx, np.int32).tobytes() + np.asarray(val, np.float32).tobytes()
    else:
        k = 0
        body = message.encode()[:4096]
    return RESP_HEAD.pack(RESP_MAGIC, RESP_HEAD.size + len(body), rid, status, k, float(latency_ms)) + body


def decode_response(head, body):
    magic, length, rid, status, k, lat = RESP_HEAD.unpack(head)
    if magic != RESP_MAGIC:
        raise ValueError("bad response frame")
    if status == 0:
        idx = np.frombuffer(body[:4 * k], np.int32).copy()
        val = np.frombuffer(body[4 * k:8 * k], np.float32).copy()
        return rid, status, idx, val, lat, ""
    return rid, status, None, None, lat, body.decode(errors="replace")


class EngineBackend:
    """The continuous-batching server over a resident database."""

    def __init__(self, db, **server_kw):
        import cudaruntime1_amd as ts
        self._ts = ts
        self.server = ts.Server(db, **server_kw)

    def submit(self, points, prm):
        kw = {f: (None if prm[f] == DEFAULT else prm[f]) for f in ("kc", "r", "band", "delta")}
        eps = None if np.isnan(prm["epsilon"]) else prm["epsilon"]
        p = self._ts.params(prm["measure"], mode=prm["mode"], k=prm["k"], epsilon=eps, **kw)
        return self.server.submit(points, p)

    def wait(self, ticket, k):
        return self.server.wait(ticket, k)

    def close(self):
        self.server.close()


class NetServer:
    """Serves `backend` (submit(points, params) -> ticket without blocking,
    wait(ticket, k) -> (idx, val, latency_ms); errors raise an exception with
    a `status`) on a TCP port from a thread of its own; up to `waiters`
    requests wait for results at once."""

    def __init__(self, backend, host="127.0.0.1", port=0, waiters=256):
        self.backend = backend
        self.host, self.port = host, port
        self._pool = ThreadPoolExecutor(max_workers=waiters)
        self._loop = None
        self._server = None
        self._ready = threading.Event()
        self._thread = None

    async def _handle(self, reader, writer):
        lock = asyncio.Lock()
        tasks = set()

        async def answer(rid, prm, pts):
            loop = asyncio.get_running_loop()
            try:
                ticket = self.backend.submit(pts, prm)  # admission answers at once
                idx, val, lat = await loop.run_in_executor(self._pool, self.backend.wait, ticket, prm["k"])
                frame = encode_response(rid, 0, idx, val, lat)
            except Exception as e:  # a refused or failed request answers with its status
                frame = encode_response(rid, int(getattr(e, "status", 1)) or 1, message=str(e))
            async with lock:
                writer.write(frame)
                await writer.drain()

        try:
            while True:
                head = await reader.readexactly(REQ_HEAD.size)
                length = struct.unpack_from("<I", head, 4)[0]
                if head[:4] != REQ_MAGIC or length < REQ_HEAD.size or length > REQ_HEAD.size + 8 * MAX_POINTS:
                    break
                body = await reader.readexactly(length - REQ_HEAD.size)
                try:
                    rid, prm, pts = decode_request(head, body)
                except ValueError:
                    break
                t = asyncio.ensure_future(answer(rid, prm, pts))
                tasks.add(t)
                t.add_done_callback(tasks.discard)
        except (asyncio.IncompleteReadError, ConnectionError):
            pass
        if tasks:
            await asyncio.gather(*tasks, return_exceptions=True)
        writer.close()

    def _run(self):
        self._loop = asyncio.new_event_loop()
        asyncio.set_event_loop(self._loop)
        self._server = self._loop.run_until_complete(asyncio.start_server(self._handle, self.host, self.port))
        self.port = self._server.sockets[0].getsockname()[1]
        self._ready.set()
        self._loop.run_forever()

    def start(self):
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()
        self._ready.wait()
        return self

    def stop(self):
        if self._loop:
            async def shutdown():
                self._server.close()
                await self._server.wait_closed()
            asyncio.run_coroutine_threadsafe(shutdown(), self._loop).result()
            self._loop.call_soon_threadsafe(self._loop.stop)
            self._thread.join()
        self._pool.shutdown(wait=True)


class RequestError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("status %d: %s" % (status, message))
        self.status = status


class Client:
    """Blocking client; `knn_many` pipelines its requests on the connection."""

    def __init__(self, host="127.0.0.1", port=7070, timeout=60.0):
        self.sock = socket.create_connection((host, port), timeout=timeout)
        self.sock.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
        self._next = 1

    def _recv(self, n):
        buf = bytearray()
        while len(buf) < n:
            chunk = self.sock.recv(n - len(buf))
            if not chunk:
                raise ConnectionError("server closed the connection")
            buf += chunk
        return bytes(buf)

    def _read_response(self):
        head = self._recv(RESP_HEAD.size)
        length = struct.unpack_from("<I", head, 4)[0]
        return decode_response(head, self._recv(length - RESP_HEAD.size))

    def knn_many(self, requests):
        """requests: [(points, params dict)]; params as knn's keywords.
        Returns [(idx, val, latency_ms)] in request order; a refused request
        gives a RequestError in its place."""
        ids = []
        frames = []
        for pts, kw in requests:
            rid = self._next
            self._next += 1
            ids.append(rid)
            frames.append(encode_request(rid, pts, **_fill(kw)))
        self.sock.sendall(b"".join(frames))
        got = {}
        while len(got) < len(ids):
            rid, status, idx, val, lat, msg = self._read_response()
            got[rid] = (idx, val, lat) if status == 0 else RequestError(status, msg)
        return [got[r] for r in ids]

    def knn(self, points, **kw):
        r = self.knn_many([(points, kw)])[0]
        if isinstance(r, Exception):
            raise r
        return r

    def close(self):
        self.sock.close()


def _fill(kw):
    """Request fields from knn keywords: measure (name or code), mode (default
    EXACT = 1), k (default 10); kc, r, band, delta and epsilon travel as
    'default' unless given."""
    measures = {"dtw": 0, "lcss": 1, "frechet": 2, "hausdorff": 3, "euclidean": 4}
    m = kw.get("measure", "dtw")
    out = dict(measure=measures[m] if isinstance(m, str) else int(m), mode=int(kw.get("mode", 1)),
               k=int(kw.get("k", 10)), epsilon=float(kw["epsilon"]) if kw.get("epsilon") is not None else float("nan"))
    for f in ("kc", "r", "band", "delta"):
        out[f] = int(kw[f]) if kw.get(f) is not None else DEFAULT
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--db", required=True, help="tsdb database file")
    ap.add_argument("--host", default="127.0.0.1")
    ap.add_argument("--port", type=int, default=7070)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args()
    import cudaruntime1_amd as ts
    ts.init(args.device)
    db = ts.TrajSet.load(args.db)
    srv = NetServer(EngineBackend(db, streams=args.streams), args.host, args.port).start()
    print("serving %d trajectories on %s:%d" % (db.count, args.host, srv.port), flush=True)
    threading.Event().wait()


if __name__ == "__main__":
    main()
