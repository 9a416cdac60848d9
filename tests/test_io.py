"""File IO keys and file naming of the reference's frame / memory-state files, on the host (no device):
GlobalParameters defaults (ParticleSolverData.h:56-64), JSON keys (ParticleSolverData.cpp:48-56),
DataIO::getFilePath (DataIO.h:119-124) and the state index loadMemoryState picks
(PIC_Solver.cpp:163-167, DataIO.h:90-98)."""
import json
import os

import banana_amd


def test_io_defaults_are_the_reference_defaults():
    io = banana_amd.io_from_json(json.dumps({"GlobalParameters": {}}))
    assert io["data_path"] == "./SimData"
    assert io["memory_state_folder"] == "MemoryState" and io["frame_data_folder"] == "FrameData"
    assert io["load_memory_state"] == 1 and io["save_memory_state"] == 0 and io["save_frame_data"] == 0
    assert io["frame_per_state"] == 1 and io["save_velocity"] == 0 and io["save_object_index"] == 0


def test_io_keys_from_json():
    scene = {"GlobalParameters": {"DataPath": "/tmp/x", "MemoryStateDataFolder": "MS", "FrameDataFolder": "FD",
                                  "LoadMemoryState": False, "SaveMemoryState": True, "SaveFrameData": True,
                                  "FramePerState": 5, "OptionalSavingData": ["ObjectIndex", "ParticleVelocity"]}}
    io = banana_amd.io_from_json(json.dumps(scene))
    assert io["data_path"] == "/tmp/x" and io["memory_state_folder"] == "MS" and io["frame_data_folder"] == "FD"
    assert (io["load_memory_state"], io["save_memory_state"], io["save_frame_data"]) == (0, 1, 1)
    assert io["frame_per_state"] == 5 and io["save_velocity"] == 1 and io["save_object_index"] == 1


def test_data_file_path_format():
    assert banana_amd.data_file_path("./SimData", "FrameData", 7) == "./SimData/FrameData/frame.0007.bnn"
    assert banana_amd.data_file_path("/a", "b", 12345) == "/a/b/frame.12345.bnn"


def test_latest_state_index(tmp_path):
    root, folder = str(tmp_path), "MemoryState"
    assert banana_amd.latest_state_index(root, folder, 1, 10) == -1
    os.makedirs(os.path.join(root, folder))
    for f in (2, 5, 9):
        open(banana_amd.data_file_path(root, folder, f), "wb").close()
    assert banana_amd.latest_state_index(root, folder, 1, 10) == 9   # highest index <= FinalFrame
    assert banana_amd.latest_state_index(root, folder, 1, 8) == 5
    assert banana_amd.latest_state_index(root, folder, 6, 10) == 5   # StartFrame - 1 exists and is > 1
    assert banana_amd.latest_state_index(root, folder, 3, 10) == 2   # StartFrame - 1 = 2 exists and is > 1
    open(banana_amd.data_file_path(root, folder, 1), "wb").close()
    assert banana_amd.latest_state_index(root, folder, 2, 10) == 9   # StartFrame - 1 = 1 is not > 1: the scan
    assert banana_amd.latest_state_index(root, folder, 1, 1) == 1
    assert banana_amd.latest_state_index(root, folder, 1, 0) == -1
