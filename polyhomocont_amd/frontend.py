"""Serving front end for clients in other processes.

Clients send request frames (wire.encode_request) over TCP. The front end
rebuilds each system with the library's own builder (add_equation,
src/sparse.f90:83-141), so a rejected equation or a non-square system gets
the same err_invalid_input as an in-process `solve`. It draws the
homotopy's constants and start points from the request's seed, as `solve`
does (src/solve.f90:166-190), and submits the request to the
continuous-batching engine (device.Server). The reply is the request's
endpoint records (wire.encode_endpoints), or an error frame. Classification
(src/solve.f90:254-426) stays with the client, on its host
(host.library().classify on the records).

One thread runs the whole loop: it accepts connections, reads frames,
submits requests as they arrive and advances the engine while any are in
flight. Replies go out in completion order, tagged with the client's
request id.

Limits keep one client from stalling the others: a request whose path
count (the product of its equations' degrees, known before anything is
built) exceeds `max_paths` -- by default the engine's ring capacity -- is
refused with a capacity error before its start points are drawn; a frame
header announcing more than `max_frame_bytes` ends that connection; any
failure while handling one request becomes that request's error reply.
"""
import math
import selectors
import socket

from . import abi, device, wire


class _Conn:
    def __init__(self, sock, max_frame_bytes):
        self.sock = sock
        self.reader = wire.FrameReader(max_frame_bytes)
        self.out = bytearray()
        self.closed = False


class FrontEnd:
    """Listens on `address` (host, port); port 0 picks a free port, read it
    back from `.address`. `server` is the engine (a device.Server, or any
    object with its submit/progress/done/release methods); `host_lib` the
    Fortran host library that builds and checks systems."""

    def __init__(self, server, host_lib, address=("127.0.0.1", 0),
                 max_paths=None, max_frame_bytes=64 << 20):
        self.srv = server
        self.H = host_lib
        if max_paths is None:
            max_paths = getattr(server, "max_request_paths", lambda: 1 << 20)()
        self.max_paths = int(max_paths)
        self.max_frame_bytes = int(max_frame_bytes)
        self.listener = socket.create_server(address)
        self.listener.setblocking(False)
        self.address = self.listener.getsockname()
        self.sel = selectors.DefaultSelector()
        self.sel.register(self.listener, selectors.EVENT_READ, None)
        self.pending = {}  # ticket -> (conn, request id, npaths, records)
        self.served = 0
        self.rejected = 0

    # -- requests -------------------------------------------------------
    def _prepare(self, usys, opts):
        """Packed system, homotopy constants and start points, or
        ValueError when the builder rejects the system."""
        g, b, a, _ = self.H.homotopy(usys, opts.seed)
        return (self.H.pack(usys), abi.HomotopyData(g, b, a), opts,
                self.H.start_points(usys, opts.seed))

    def _reply(self, conn, frame):
        if not conn.closed:
            conn.out += frame
            self._flush(conn)

    def _request(self, conn, frame):
        try:
            rid, usys, opts = wire.decode_request(frame)
        except wire.WireError as e:
            self.rejected += 1
            self._reply(conn, wire.encode_error(0, wire.PHC_ERR_INVALID_INPUT,
                                                str(e)))
            return
        try:
            degrees = usys.degrees
            paths = math.prod(degrees) if degrees and min(degrees) > 0 else 0
            if paths > self.max_paths:
                self.rejected += 1
                self._reply(conn, wire.encode_error(
                    rid, wire.PHC_ERR_CAPACITY,
                    f"{paths} paths exceed the server's {self.max_paths}"))
                return
            req = self._prepare(usys, opts)
            ticket, out = self.srv.submit(*req)
        except ValueError as e:
            self.rejected += 1
            self._reply(conn, wire.encode_error(
                rid, wire.PHC_ERR_INVALID_INPUT, f"invalid system: {e}"))
            return
        except device.DeviceError as e:
            self.rejected += 1
            self._reply(conn, wire.encode_error(rid, e.code or -1, str(e)))
            return
        except Exception as e:  # noqa: BLE001 -- one request, not the loop
            self.rejected += 1
            self._reply(conn, wire.encode_error(
                rid, -1, f"request failed: {type(e).__name__}: {e}"))
            return
        self.pending[ticket] = (conn, rid, len(req[3]), out)

    def _finish(self):
        for t in [t for t in self.pending if self.srv.done(t)]:
            conn, rid, npaths, out = self.pending.pop(t)
            self.srv.release(t)
            self.served += 1
            self._reply(conn, wire.encode_endpoints(rid, out, npaths))

    # -- sockets --------------------------------------------------------
    def _flush(self, conn):
        if conn.closed:
            return
        try:
            while conn.out:
                n = conn.sock.send(conn.out)
                del conn.out[:n]
        except BlockingIOError:
            pass
        except OSError:
            self._close(conn)
            return
        events = selectors.EVENT_READ | (selectors.EVENT_WRITE if conn.out
                                         else 0)
        self.sel.modify(conn.sock, events, conn)

    def _close(self, conn):
        if conn.closed:
            return
        conn.closed = True
        self.sel.unregister(conn.sock)
        conn.sock.close()

    def _readable(self, conn):
        try:
            data = conn.sock.recv(1 << 20)
        except BlockingIOError:
            return
        except OSError:
            data = b""
        if not data:
            self._close(conn)
            return
        try:
            frames = conn.reader.feed(data)
        except wire.WireError as e:
            # the stream cannot be resynchronised: answer once and drop it
            self.rejected += 1
            conn.out += wire.encode_error(0, wire.PHC_ERR_INVALID_INPUT,
                                          str(e))
            self._flush(conn)
            self._close(conn)
            return
        for f in frames:
            self._request(conn, f)

    def step(self, timeout=0.05):
        """One turn of the loop: socket events (waiting at most `timeout`
        seconds when nothing is in flight), then one engine round and the
        replies of finished requests."""
        for key, events in self.sel.select(0 if self.pending else timeout):
            if key.data is None:
                try:
                    s, _ = self.listener.accept()
                except BlockingIOError:
                    continue
                s.setblocking(False)
                s.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
                c = _Conn(s, self.max_frame_bytes)
                self.sel.register(s, selectors.EVENT_READ, c)
                continue
            conn = key.data
            if events & selectors.EVENT_WRITE:
                self._flush(conn)
            if events & selectors.EVENT_READ and not conn.closed:
                self._readable(conn)
        if self.pending:
            self.srv.progress()
            self._finish()

    def serve(self, stop):
        """Runs until `stop` (a threading.Event) is set and nothing is in
        flight."""
        while not stop.is_set() or self.pending:
            self.step()

    def close(self):
        for key in list(self.sel.get_map().values()):
            if key.data is not None:
                self._close(key.data)
        self.sel.unregister(self.listener)
        self.listener.close()
        self.sel.close()


class Client:
    """Blocking client of a FrontEnd."""

    def __init__(self, address, timeout=120.0):
        self.sock = socket.create_connection(address, timeout=timeout)
        self.sock.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)

    def send(self, usys, opts, request_id):
        wire.send_frame(self.sock, wire.encode_request(usys, opts,
                                                       request_id))

    def receive(self):
        """(request id, EndpointArrays or None, error code, message) of the
        next reply, or None when the server closed the stream."""
        f = wire.recv_frame(self.sock)
        return None if f is None else wire.decode_endpoints(f)

    def close(self):
        self.sock.close()
