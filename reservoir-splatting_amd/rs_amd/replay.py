"""Camera record/replay of the reference's modified Falcor Scene, which drives the pass's
frozen mode and the accumulation flags (Source/Falcor/Scene/Scene.cpp:1567-1590,
2403-2486).  A recorded camera sequence is replayed frame by frame:

* the first frame of every pass through the sequence sets Scene::shouldReset, so the pass
  drops its temporal history and refills its camera keyframes (:2424-2432);
* without "Replay for Accumulation" the scene freezes once the sequence reaches the
  accumulation frame (Scene::shouldFreeze, :2459-2470), and a frozen scene stops
  replaying: every later frame re-renders that frame from the same history;
* with "Replay for Accumulation" the sequence loops and never freezes, and only the
  accumulation frame is accumulated (Scene::shouldAccumulate, :2474-2479).

`step()` returns what the pass reads from the Scene for the next frame; `execute()`
renders it and returns the pass's dictionary entries for the AccumulatePass."""
from . import abi


class CameraReplay:
    def __init__(self, cameras, accumulation_frame=-1, replay_accumulation=False, view_accumulation=False):
        self.cameras = list(cameras)
        if not self.cameras:
            raise ValueError("an empty camera sequence")
        self.accumulation_frame = int(accumulation_frame)
        self.replay_accumulation = bool(replay_accumulation)
        self.view_accumulation = bool(view_accumulation)
        self.recorded_frame = 0   # mRecordedFrameCount: the next frame of the sequence
        self.frozen = False       # mFrozen
        self.should_reset = False  # mShouldReset
        self.camera = self.cameras[0]

    def _current(self):
        n = len(self.cameras)
        return (self.recorded_frame - 1 + n) % n

    def step(self):
        """(camera, rs_frame_io.sceneFlags, frozen) of the next frame."""
        if not self.frozen:
            # Scene::replayStep
            frame = self.recorded_frame
            self.should_reset = frame == 0
            self.camera = self.cameras[frame]
            self.recorded_frame = (self.recorded_frame + 1) % len(self.cameras)
            self.frozen = (not self.replay_accumulation) and self._current() == self.accumulation_frame
        flags = abi.SCENE_REPLAYING
        if self.should_reset:
            flags |= abi.SCENE_SHOULD_RESET
        if self._current() == self.accumulation_frame:
            flags |= abi.SCENE_ACCUMULATION_FRAME
        if self.view_accumulation and self.accumulation_frame >= 0:
            flags |= abi.SCENE_VIEW_ACCUMULATION
        return self.camera, flags, self.frozen

    def execute(self, rs, width, height, **kw):
        """Renders the next frame of the replay with `rs` (a ReservoirSplatting) and returns
        its dictionary entries (_suppressAutoReset, _shouldAccumulate, _viewAccumulation)."""
        cam, flags, frozen = self.step()
        if rs.frozen != frozen:
            rs.frozen = frozen
        rs.execute(cam, width, height, sceneFlags=flags, **kw)
        return rs.frameFlags()
